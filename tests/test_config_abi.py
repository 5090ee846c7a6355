"""CPU tests of the boundary: the library loads and exports every function
include/opensayal.h declares; the config loader keeps the reference's key
surface, defaults and get_or lookup rules (inc/config_parser.hpp:131-150,
src/config_parser.cpp:21-181); invalid configurations are rejected before
any device work."""
import ctypes
import json
import re
from pathlib import Path

import pytest

from opensayal_amd import ConfigParser, OpenSayalError, _abi, default_config
from tests.helpers import make_config

ROOT = Path(__file__).resolve().parents[1]
GOLDEN = Path(__file__).resolve().parent / "golden"


def declared_functions():
    text = (ROOT / "include" / "opensayal.h").read_text()
    return sorted(set(re.findall(r"^\s*(?:int|const char\*|void)\s+(osy_\w+)\s*\(", text, re.M)))


def test_library_exports_every_declared_function():
    lib = _abi.lib()
    names = declared_functions()
    assert len(names) >= 18
    for name in names:
        assert hasattr(lib, name), name
        assert name in _abi.PROTOTYPES, name
    assert lib.osy_abi_version() == 7


def test_struct_sizes_match_header_layout():
    # 34 four-byte fields, no padding
    assert ctypes.sizeof(_abi.OsyConfig) == 34 * 4
    assert ctypes.sizeof(_abi.OsySource) == 5 * 4
    assert ctypes.sizeof(_abi.OsyBatchTuning) == 3 * 4
    # count, groups, then four 8-byte fields
    assert ctypes.sizeof(_abi.OsyBatchInfo) == 8 + 4 * 8
    # four 4-byte fields, then four 8-byte fields
    assert ctypes.sizeof(_abi.OsyBatchItemTime) == 4 * 4 + 4 * 8


DOCUMENTED_DEFAULTS = dict(
    width=1920, height=1080, cell_size=1.0, cell_pixel_size=1, enable_drain=1,
    enable_pressure=0, enable_smoke=1, enable_interactive=0, projection_n=50,
    projection_o=1.9, pipe_height=270, pipe_length=0, smoke_length=1, smoke_height=270,
    smoke_count=1, wind_speed=0.0, wind_smoke=1.0, g=0.0, d_t=0.05, enable_real_time=0,
    real_time_multiplier=1.0, enable_decay=0, decay_rate=0.05, obstacle_enable=1,
    obstacle_center_x=960, obstacle_center_y=540, obstacle_radius=36.0, density=1.0,
    drag_coeff=0.0, block_size_x=64, block_size_y=1,
)


def test_empty_config_gives_documented_defaults():
    c = ConfigParser.parse_string("{}")
    for k, v in DOCUMENTED_DEFAULTS.items():
        assert getattr(c, k) == pytest.approx(v), k
    assert c.viscosity == pytest.approx(0.001)  # fluid.viscosity default


def test_size_dependent_defaults_follow_the_grid():
    c = ConfigParser.parse_string('{"sim": {"width": 300, "height": 200}}')
    assert (c.pipe_height, c.smoke_height) == (50, 50)
    assert (c.obstacle_center_x, c.obstacle_center_y) == (150, 100)
    assert c.obstacle_radius == pytest.approx(200 / 30.0)


@pytest.mark.parametrize("text", [
    '{"sim": {"time": {"d_t": 0.125}}}',
    '{"sim.time.d_t": 0.125}',
    '{"sim": {"time.d_t": 0.125}}',
])
def test_nested_and_dotted_keys_are_equivalent(text):
    assert ConfigParser.parse_string(text).d_t == 0.125


def test_direct_hit_wins_over_nested():
    c = ConfigParser.parse_string('{"sim": {"width": 100}, "sim.width": 200}')
    assert c.width == 200


def test_type_errors_top_level_fail_nested_default():
    with pytest.raises(OpenSayalError) as e:
        ConfigParser.parse_string('{"sim.width": "wide"}')
    assert e.value.status == _abi.OSY_E_CONFIG
    c = ConfigParser.parse_string('{"sim": {"width": "wide", "height": 90}}')
    assert c.width == 1920 and c.height == 90
    c = ConfigParser.parse_string('{"sim": {"enable_drain": 0}}')  # bool needs a boolean
    assert c.enable_drain == 1
    c = ConfigParser.parse_string('{"sim": {"width": 640.9, "projection": {"n": true}}}')
    assert c.width == 640 and c.projection_n == 1  # numbers truncate, booleans count


def test_malformed_or_missing_file(tmp_path):
    p = tmp_path / "bad.json"
    p.write_text('{"sim": {"width": 10,}}')
    with pytest.raises(OpenSayalError) as e:
        ConfigParser(str(p)).parse()
    assert e.value.status == _abi.OSY_E_CONFIG
    with pytest.raises(OpenSayalError):
        ConfigParser(str(tmp_path / "absent.json")).parse()


def test_full_example_file():
    """A complete OpenSayal.conf.json of the documented shape (render keys
    included, which are accepted and have no effect)."""
    c = ConfigParser(str(GOLDEN / "OpenSayal.conf.json")).parse()
    exp = json.loads((GOLDEN / "OpenSayal.conf.expected.json").read_text())
    for k, v in exp.items():
        assert getattr(c, k) == pytest.approx(v), k


def test_storage_key():
    assert ConfigParser.parse_string('{"sim": {"storage": "bf16"}}').storage == 2
    with pytest.raises(OpenSayalError):
        ConfigParser.parse_string('{"sim": {"storage": "f8"}}')


@pytest.mark.parametrize("bad,status", [
    (dict(viscosity=float("nan")), _abi.OSY_E_INVALID),
    (dict(block_size_x=64, obstacle_radius=2.0e7), _abi.OSY_E_INVALID),
    (dict(projection_o=2.5), _abi.OSY_E_INVALID),
    (dict(projection_n=-1), _abi.OSY_E_INVALID),
    (dict(cell_size=0.5), _abi.OSY_E_INVALID),
    (dict(smoke_count=2, pipe_height=10, smoke_height=10), _abi.OSY_E_INVALID),
])
def test_create_rejects_invalid_configs_before_device_work(bad, status):
    c = make_config(64, 36, **bad)
    h = ctypes.c_void_p()
    st = _abi.lib().osy_fluid_create(ctypes.byref(c), ctypes.byref(h))
    assert st == status
    assert _abi.lib().osy_last_error()


def test_default_config_helper_sets_overrides():
    c = default_config(64, 36, wind_speed=2.0, storage="f16")
    assert c.width == 64 and c.wind_speed == 2.0 and c.storage == 1 and c.viscosity == 0.0


def test_diagnostic_entry_points_reject_bad_arguments_before_device_work():
    """The timing, per-grid threshold and arithmetic-check entry points
    validate their arguments before touching a device."""
    lib = _abi.lib()
    n = ctypes.c_uint64()
    first = ctypes.c_uint32()
    assert lib.osy_arith_check(8, ctypes.byref(n), ctypes.byref(first)) == _abi.OSY_E_INVALID
    assert b"mode" in lib.osy_last_error()
    assert lib.osy_arith_check(-1, ctypes.byref(n), ctypes.byref(first)) == _abi.OSY_E_INVALID
    assert lib.osy_arith_check(0, None, ctypes.byref(first)) == _abi.OSY_E_INVALID
    assert b"null" in lib.osy_last_error()
    count = ctypes.c_int64()
    assert lib.osy_batch_set_timing(None, 1) == _abi.OSY_E_INVALID
    assert lib.osy_batch_set_per_grid_cells(None, 0) == _abi.OSY_E_INVALID
    assert lib.osy_batch_item_times(None, None, 0, ctypes.byref(count), None) == \
        _abi.OSY_E_INVALID
