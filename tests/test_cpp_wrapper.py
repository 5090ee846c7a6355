"""The C++ wrapper with the reference's surface (include/opensayal.hpp)
compiles on its own and links against the library: a program written the way
src/main.cu drives Fluid (ConfigParser, Fluid(Config), update(Source, d_t),
min/max_pressure, width/height) builds unchanged apart from the include.
Running it needs a GPU; the headless program, built on the same wrapper, is
exercised by tests/test_gpu_cli.py."""
import shutil
import subprocess
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
CXX = shutil.which("g++")

PROGRAM = r'''
#include "opensayal.hpp"
#include <cstdio>
int main() {
  auto config = opensayal::ConfigParser("OpenSayal.conf.json").parse();
  opensayal::Fluid fluid(config);
  opensayal::Source source;
  source.active = true;
  source.smoke = 1.0f;
  source.velocity = 200.0f;
  source.position = opensayal::Vector2d<int>(fluid.width / 2, fluid.height / 2);
  fluid.update(source, config.d_t);
  std::vector<float> u = fluid.field(OSY_FIELD_VEL_X);
  std::vector<uint32_t> frame = fluid.render_rgba();
  std::printf("%d %d %g %g %zu %zu\n", fluid.width, fluid.height, fluid.min_pressure,
              fluid.max_pressure, u.size(), frame.size());
  return 0;
}
'''


@pytest.mark.skipif(CXX is None, reason="no host C++ compiler")
def test_wrapper_compiles_and_links(tmp_path):
    lib_dir = ROOT / "opensayal_amd"
    if not (lib_dir / "libopensayal_amd.so").exists():
        pytest.fail("libopensayal_amd.so is missing: run `python -m opensayal_amd.build`")
    src = tmp_path / "app.cpp"
    src.write_text(PROGRAM)
    out = tmp_path / "app"
    cmd = [CXX, "-std=c++17", "-Wall", "-Wextra", "-Werror", f"-I{ROOT / 'include'}", str(src),
           f"-L{lib_dir}", "-lopensayal_amd", f"-Wl,-rpath,{lib_dir}", "-o", str(out)]
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    assert out.exists()


@pytest.mark.skipif(CXX is None, reason="no host C++ compiler")
def test_wrapper_reports_errors_as_exceptions(tmp_path):
    """A missing config file surfaces as opensayal::Error with the library's
    message and status, before any device work."""
    src = tmp_path / "err.cpp"
    src.write_text(r'''
#include "opensayal.hpp"
#include <cstdio>
int main() {
  try {
    opensayal::ConfigParser("/nonexistent/OpenSayal.conf.json").parse();
  } catch (const opensayal::Error& e) {
    std::printf("%d %s\n", e.status(), e.what());
    return 0;
  }
  return 1;
}
''')
    lib_dir = ROOT / "opensayal_amd"
    out = tmp_path / "err"
    subprocess.run([CXX, "-std=c++17", f"-I{ROOT / 'include'}", str(src), f"-L{lib_dir}",
                    "-lopensayal_amd", f"-Wl,-rpath,{lib_dir}", "-o", str(out)], check=True)
    r = subprocess.run([str(out)], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0
    assert "cannot open config file" in r.stdout
    assert int(r.stdout.split()[0]) == -5  # OSY_E_CONFIG
