"""The headless fluid-simulator program: argument and config errors end
with status 1 and a message before anything touches the GPU (the reference's
usage check, src/main.cu:18-25)."""
import json
import subprocess

import pytest

from opensayal_amd._abi import CLI_PATH


def run(*args, cwd=None):
    return subprocess.run([str(CLI_PATH), *args], capture_output=True, text=True, timeout=60,
                          cwd=cwd)


@pytest.fixture(scope="module", autouse=True)
def built():
    if not CLI_PATH.exists():
        pytest.fail(f"{CLI_PATH} is missing: run `python -m opensayal_amd.build`")


def test_help():
    r = run("--help")
    assert r.returncode == 0 and "usage: fluid-simulator" in r.stdout


@pytest.mark.parametrize("args,msg", [
    (["--bogus"], "unknown argument --bogus"),
    (["--steps"], "--steps needs N >= 0"),
    (["--steps", "-3"], "--steps needs N >= 0"),
    (["--steps", "12x"], "--steps needs N >= 0"),
    (["--save-report", "--compare-performance"], "exclusive"),
    (["--dump-every", "5"], "--dump-every needs --dump"),
    (["--frame"], "--frame and --traces write into the --dump directory"),
    (["--dump-every", "0", "--dump", "d"], "--dump-every needs K >= 1"),
])
def test_usage_errors(args, msg):
    r = run(*args)
    assert r.returncode == 1
    assert msg in r.stderr and "usage: fluid-simulator" in r.stderr


def test_missing_default_config(tmp_path):
    r = run("--steps", "1", cwd=tmp_path)  # no OpenSayal.conf.json here
    assert r.returncode == 1 and "cannot open config file OpenSayal.conf.json" in r.stderr


@pytest.mark.parametrize("text,msg", [
    ("{not json", "config"),
    (json.dumps({"sim": {"storage": "f64"}}), "sim.storage"),
    # a direct hit on the whole dotted key converts or fails (get_or)
    (json.dumps({"sim.width": "wide"}), "config"),
    # a non-finite viscosity is refused when the solver is created
    (json.dumps({"sim": {"width": 64, "height": 48}, "fluid": {"viscosity": 1e39}}),
     "fluid.viscosity must be finite"),
])
def test_config_errors(tmp_path, text, msg):
    cfg = tmp_path / "c.json"
    cfg.write_text(text)
    r = run("--config", str(cfg), "--steps", "1")
    assert r.returncode == 1 and msg in r.stderr, r.stderr


def test_bad_path_line_settings(tmp_path):
    cfg = tmp_path / "c.json"
    cfg.write_text(json.dumps({"visual": {"path_line": {"enable": True, "distance": 0}}}))
    r = run("--config", str(cfg), "--steps", "1")
    assert r.returncode == 1 and "visual.path_line" in r.stderr
