"""The allocation-free residual sweep (xb_elast_residual_noalloc: squared-norm
partials in the library's static device array) against the residual sweep with a
caller workspace (xb_elast_step in residual mode, then xb_norm_finalize): the
restricted residual and the (norm, non-finite flag) pair bit for bit, on two calls
in a row, the second on a smaller grid (the static array keeps the first call's
contents), and a sweep larger than the array rejected before any launch."""
import pytest
import torch

import problems as P
from src import _native as nat

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)


def _both(el, level, f, resid, ext):
    """(coarse, norm) of the workspace path and of the allocation-free path."""
    coarse_a, ws, n = el._residual(f, level, resid, ext)
    norm_a = ws.finalize(n)
    coarse_b = torch.full_like(coarse_a, float("nan"))
    a = el._launch(f, None, level, nat.XB_MODE_RESIDUAL, ext=ext, resid=resid, coarse=coarse_b)
    a.norm_partials = None
    norm_b = torch.full((2,), float("nan"), dtype=torch.float64, device=f.device)
    nat.check(nat.lib().xb_elast_residual_noalloc(a, nat.ptr(norm_b), nat.stream()), "xb_elast_residual_noalloc")
    torch.cuda.synchronize()
    return coarse_a, norm_a, coarse_b, norm_b, n


def test_noalloc_residual_matches_workspace_path_twice():
    g = torch.Generator().manual_seed(11)
    # first call: 256^2 periodic grid, finest level, body force + external force
    el = P.periodic_elasticity("gpu", nx=256, levels=2)
    f = (1e-3 * torch.randn((8, 256, 256), generator=g, dtype=torch.float64)).cuda()
    ext = (1e-4 * torch.randn((2, 256, 256), generator=g, dtype=torch.float64)).cuda()
    ca, na, cb, nb, n = _both(el, 0, f, None, ext)
    say("256^2 level 0:", n, "partials, norm", na[0].item(), nb[0].item())
    assert n > 1
    assert torch.equal(ca, cb)
    assert torch.equal(na, nb)
    assert na[1].item() == 0.0
    # second call, smaller: traction walls, level 1 of a 3-level hierarchy, with a
    # restricted residual input subtracted on the interior box
    el2 = P.neumann_elasticity("gpu", n=32, levels=3)
    nx, ny = el2.grid_size(1)
    x0, x1, y0, y1 = el2.interior_box(1)
    f2 = (1e-3 * torch.randn((8, nx, ny), generator=g, dtype=torch.float64)).cuda()
    r2 = (1e-4 * torch.randn((8, x1 - x0, y1 - y0), generator=g, dtype=torch.float64)).cuda()
    ca, na, cb, nb, n2 = _both(el2, 1, f2, r2, None)
    say(f"{nx}x{ny} level 1:", n2, "partials, norm", na[0].item(), nb[0].item())
    assert n2 < n
    assert torch.equal(ca, cb)
    assert torch.equal(na, nb)


def test_noalloc_residual_rejects_more_partials_than_the_scratch_holds():
    # 4096^2 with two rows per wavefront: 67 strips x 2048 row blocks = 137216 items
    el = P.periodic_elasticity("gpu", nx=4096, levels=2)
    f = torch.zeros((8, 4096, 4096), dtype=torch.float64, device="cuda")
    x0, x1, y0, y1 = el.interior_box(0)
    coarse = torch.zeros((8, (x1 - x0) // 2, (y1 - y0) // 2), dtype=torch.float64, device="cuda")
    a = el._launch(f, None, 0, nat.XB_MODE_RESIDUAL, coarse=coarse)
    a.rows = 2
    n = int(nat.lib().xb_elast_partials_count(a))
    say("4096^2, rows 2:", n, "partials; static scratch", nat.STATIC_PARTIALS)
    assert n > nat.STATIC_PARTIALS
    a.norm_partials = None
    out = torch.zeros(2, dtype=torch.float64, device="cuda")
    with pytest.raises(ValueError, match="static scratch holds"):
        nat.check(nat.lib().xb_elast_residual_noalloc(a, nat.ptr(out), nat.stream()), "xb_elast_residual_noalloc")
    # nothing was launched: the output is untouched
    torch.cuda.synchronize()
    assert torch.equal(out, torch.zeros_like(out))
