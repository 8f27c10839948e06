"""The launch-width table behind the autotuner (CPU: no device call)."""
import json

import pytest

from vmc_torch_amd import _lib, autotune


def test_width_table_set_get_and_reset():
    try:
        assert autotune.get_waves("sweep", 64, False) == 0
        autotune.set_waves("sweep", 64, False, 2)
        autotune.set_waves("local_energy", 32, True, 8)
        assert autotune.get_waves("sweep", 64, False) == 2
        assert autotune.get_waves("local_energy", 32, True) == 8
        # other keys and kinds untouched
        assert autotune.get_waves("sweep", 64, True) == 0
        assert autotune.get_waves("refresh", 64, False) == 0
    finally:
        autotune.set_waves("sweep", 64, False, 0)
        autotune.set_waves("local_energy", 32, True, 0)
    assert autotune.get_waves("sweep", 64, False) == 0


@pytest.mark.parametrize("waves", [3, 16, -1])
def test_width_table_rejects_bad_widths(waves):
    with pytest.raises(_lib.VmcxError, match="waves"):
        autotune.set_waves("refresh", 16, False, waves)


def test_width_table_rejects_unknown_kind():
    lib = _lib.lib()
    assert lib.vmcx_set_launch_waves(7, 0, 2) == 1
    assert "kind" in lib.vmcx_last_error().decode()
    with pytest.raises(KeyError):
        autotune.set_waves("gram", 16, False, 2)


def test_missing_cache_installs_nothing(tmp_path):
    assert autotune.load_cache(str(tmp_path / "none.json")) == 0
    bad = tmp_path / "bad.json"
    bad.write_text(json.dumps([1, 2]))
    with pytest.raises(ValueError):
        autotune.load_cache(str(bad))
