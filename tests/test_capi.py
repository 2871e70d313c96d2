"""The C-ABI library builds, loads without a GPU and exports exactly what
``include/xlb_biot.h`` declares; host-side validation of the solver API."""
import ctypes
import os
import re

import numpy as np
import pytest

from src import _native as nat

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    text = open(os.path.join(ROOT, "include", "xlb_biot.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"^\s*(?:const\s+)?[a-z_0-9]+\s*\**\s*(xb_[a-z_0-9]+)\s*\(", text, flags=re.M)))


def test_header_declares_the_entry_points():
    names = declared_functions()
    for n in ("xb_elast_step", "xb_darcy_step", "xb_darcy_gradient", "xb_elast_prolong_add", "xb_norm_finalize"):
        assert n in names


def test_library_exports_every_declared_symbol():
    lib = nat.load_library()
    for n in declared_functions():
        assert hasattr(lib, n), n
    assert set(declared_functions()) == set(nat.SIGNATURES), "ctypes bindings and header disagree"
    assert lib.xb_abi_version() == 1


def _integration_stub():
    """Namespace left by the reference-side ctypes binding shown in INTEGRATION.md."""
    text = open(os.path.join(ROOT, "INTEGRATION.md")).read()
    section = text.split("## 3. The reference-side binding", 1)[1].split("\n## ", 1)[0]
    code = re.search(r"```python\n(.*?)```", section, flags=re.S).group(1)
    ns = {}
    cwd = os.getcwd()
    os.chdir(ROOT)  # the stub loads the library by its path from the repository root
    try:
        exec(compile(code, "INTEGRATION.md", "exec"), ns)
    finally:
        os.chdir(cwd)
    return ns


def test_integration_stub_declares_the_whole_argument_struct():
    # the documented binding must lay out xb_elast_args exactly as the library reads it
    stub = _integration_stub()["ElastArgs"]
    assert ctypes.sizeof(stub) == ctypes.sizeof(nat.ElastArgs)
    got = [(n, getattr(stub, n).offset, getattr(stub, n).size) for n, _ in stub._fields_]
    want = [(n, getattr(nat.ElastArgs, n).offset, getattr(nat.ElastArgs, n).size) for n, _ in nat.ElastArgs._fields_]
    assert got == want


STRUCTS = (("xb_elast_args", nat.ElastArgs), ("xb_darcy_args", nat.DarcyArgs), ("xb_bc_table", nat.BCTable),
           ("xb_prolong_job", nat.ProlongJob), ("xb_fill_job", nat.FillJob), ("xb_norm_job", nat.NormJob),
           ("xb_batch_plan", nat.BatchPlan), ("xb_converge_state", nat.ConvergeState),
           ("xb_elast_adjoint_args", nat.ElastAdjointArgs), ("xb_darcy_adjoint_args", nat.DarcyAdjointArgs),
           ("xb_elast_mg_adjoint_args", nat.ElastMGAdjointArgs), ("xb_tail_level", nat.TailLevel),
           ("xb_tail_plan", nat.TailPlan))


def test_struct_sizes_and_offsets_match_the_c_compiler(tmp_path):
    # sizeof / offsetof of every struct as gcc lays it out, against the ctypes mirrors
    import shutil
    import subprocess
    if shutil.which("gcc") is None:
        pytest.skip("gcc not available")
    lines = ['#include <stdio.h>', '#include <stddef.h>', '#include "xlb_biot.h"', "int main(void) {"]
    for struct, cls in STRUCTS:
        lines.append('printf("%s %%zu\\n", sizeof(%s));' % (struct, struct))
        for name, _ in cls._fields_:
            lines.append('printf("%s.%s %%zu\\n", offsetof(%s, %s));' % (struct, name, struct, name))
    lines.append("return 0; }")
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines))
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)], check=True)
    out = dict(l.split() for l in subprocess.run([str(exe)], check=True, capture_output=True,
                                                  text=True).stdout.splitlines())
    for struct, cls in STRUCTS:
        assert int(out[struct]) == ctypes.sizeof(cls), struct
        for name, _ in cls._fields_:
            assert int(out[f"{struct}.{name}"]) == getattr(cls, name).offset, (struct, name)


def test_struct_layouts_match_header():
    # the ctypes mirrors of the argument structs follow the header field by field
    text = open(os.path.join(ROOT, "include", "xlb_biot.h")).read()
    for struct, cls in STRUCTS:
        body = re.search(r"typedef struct %s \{(.*?)\} %s;" % (struct, struct), text, flags=re.S).group(1)
        body = re.sub(r"/\*.*?\*/", "", body, flags=re.S)
        fields = []
        for decl in body.split(";"):
            decl = decl.strip()
            if not decl:
                continue
            names = decl.split()[-1] if "," not in decl else decl
            for part in re.split(r",", decl):
                fields.append(re.findall(r"[A-Za-z_][A-Za-z_0-9]*", part)[-1])
        assert fields == [f[0] for f in cls._fields_], (struct, fields)


def test_solver_validation_errors_match_reference():
    import problems as P
    from src.elasticity import LBMElasticity
    from src.lattice import LatticeD2Q8, LatticeD2Q9
    kw = {"lambda": 1.0, "mu": 1.0, "nx": 16, "ny": 16, "nz": 0, "precision": "f64/f64",
          "lattice": LatticeD2Q8("f64/f64"), "print_info_rate": 0}
    LBMElasticity(**kw)
    with pytest.raises(TypeError):
        LBMElasticity(**{**kw, "lambda": 1})
    with pytest.raises(ValueError):
        LBMElasticity(**{**kw, "MG_levels": 6})
    with pytest.raises(ValueError):
        LBMElasticity(**{**kw, "MG_relaxation": 1.5})
    with pytest.raises(TypeError):
        LBMElasticity(**{**kw, "MG_BCs": {"left": 1}})
    with pytest.raises(ValueError):
        LBMElasticity(**{**kw, "lattice": LatticeD2Q9("f64/f64")})
    # elasticity.py:606-629: "f64/f32", "f64/f16" and "f32/f16" are arithmetic in the
    # first precision with storage in the second; FP16 arithmetic is not carried
    from src import _native as nat
    for prec, cdt, odt, code in (("f64/f32", np.float64, np.float32, nat.XB_F64_S32),
                                 ("f64/f16", np.float64, np.float16, nat.XB_F64_S16),
                                 ("f32/f16", np.float32, np.float16, nat.XB_F32_S16)):
        pp = LBMElasticity(**{**kw, "precision": prec}).precisionPolicy
        assert pp.mixed and pp.compute_dtype is cdt and pp.output_dtype is odt and pp.storage_code == code
    with pytest.raises(NotImplementedError):
        LBMElasticity(**{**kw, "precision": "f16/f16"})
    with pytest.raises(TypeError):
        P.biot_periodic("gpu", nx=16, coupling="sideways")


def test_kernels_refuse_to_run_without_gpu():
    import torch
    if torch.cuda.is_available():
        pytest.skip("a GPU is visible")
    import problems as P
    el = P.periodic_elasticity("gpu", nx=16)
    f = el.assign_fields_sharded()
    with pytest.raises(RuntimeError, match="ROCm GPU"):
        el.step(f, 0)
