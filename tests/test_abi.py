"""C-ABI library checks that need no GPU: it loads, exports every symbol include/homo3d.h declares,
and its host-only entry point (the seeded density generator) agrees with the oracle."""
import os
import re

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    src = open(os.path.join(ROOT, "include", "homo3d.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(homo3d_[a-z0-9_]+)\s*\(", src)))


def test_library_exports_every_declared_symbol():
    import homo3d_amd

    L = homo3d_amd.lib()
    names = declared_symbols()
    assert len(names) >= 20
    missing = [s for s in names if not hasattr(L, s)]
    assert not missing, missing
    # and the Python binding wires every declared symbol
    assert set(names) <= set(homo3d_amd.api.EXPORTS)


@pytest.mark.parametrize("n,seed,sym", [(16, 0, "reflect6"), (24, 5, "rotate3"), (20, 2, "reflect3"), (12, 1, "none")])
def test_randc_density_matches_oracle(orc, n, seed, sym):
    import homo3d_amd

    a = homo3d_amd.randc_density(n, seed, 0.3, sym)
    b = orc.randc_density(n, seed, 0.3, sym)
    assert np.abs(a - b).max() <= 1e-6
    assert abs(a.astype(np.float64).mean() - 0.3) < 1e-3
    assert a.min() >= 1e-3 and a.max() <= 1.0
    # the projection is exact: the product's field is invariant under the group (bit-exact)
    if sym != "none":
        assert np.array_equal(orc.symmetrize(n, a, sym), a)


def test_bad_config_is_rejected_without_gpu():
    import homo3d_amd

    with pytest.raises(homo3d_amd.Homo3dError) as ei:
        homo3d_amd.Homogenization(homo3d_amd.HomoConfig(reso=[32, 32, 16]))
    assert ei.value.code == homo3d_amd.HOMO3D_NO_SUPPORT
    with pytest.raises(homo3d_amd.Homo3dError):
        homo3d_amd.Homogenization(homo3d_amd.HomoConfig(reso=4))


@pytest.mark.parametrize("n", [320, 384, 448, 512, 1024])
def test_sizes_beyond_32bit_offsets_are_rejected_without_gpu(n):
    """The FP64 iterate [6][3][n^3] must stay below 2^32 bytes for the gathers' 32-bit byte offsets:
    144 n^3 < 2^32 fails for every solvable n > 256 (the next after 256 is 320 = 5 * 2^6)."""
    import homo3d_amd

    assert 144 * n ** 3 >= 2 ** 32
    for solver in ("pcg", "mg"):
        with pytest.raises(homo3d_amd.Homo3dError) as ei:
            homo3d_amd.Homogenization(homo3d_amd.HomoConfig(reso=n, solver=solver))
        assert ei.value.code == homo3d_amd.HOMO3D_NO_SUPPORT
    assert 144 * 256 ** 3 < 2 ** 32
