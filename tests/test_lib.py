"""CPU checks of the C-ABI library: it loads, exports every symbol the
header declares, and rejects bad arguments before touching a device."""
import ctypes

import pytest

from vmc_torch_amd import _lib


def test_library_exports_header_symbols():
    lib = _lib.lib()
    names = _lib.header_functions()
    assert len(names) >= 8
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    assert set(_lib._SIGNATURES) == set(names)


def test_abi_version_and_nfp():
    lib = _lib.lib()
    assert lib.vmcx_abi_version() == 19
    assert [lib.vmcx_slater_nfp(n) for n in (1, 8, 9, 16, 17, 32, 33, 64, 65)] == \
        [8, 8, 16, 16, 32, 32, 64, 64, -1]


@pytest.mark.parametrize("n_sites,n_f,msg", [(100, 8, "n_sites"), (8, 65, "n_f"), (4, 9, "n_f")])
def test_argument_errors_are_reported(n_sites, n_f, msg):
    lib = _lib.lib()
    rc = lib.vmcx_slater_refresh(1, n_sites, n_f, None, None, None, None, None, None,
                                 None, -1, None)
    assert rc == 1
    assert msg in lib.vmcx_last_error().decode()
    with pytest.raises(_lib.VmcxError):
        _lib.check(rc)


def test_sweep_rejects_partial_injected_draws():
    lib = _lib.lib()
    p = ctypes.c_void_p(16)  # never dereferenced: the check fails first
    rc = lib.vmcx_slater_sweep(1, 4, 2, p, 1, p, 0.25, 1, 0, 0, p, None, None, 0,
                               p, p, p, p, p, p, None, None, None)
    assert rc == 1 and "injected" in lib.vmcx_last_error().decode()
