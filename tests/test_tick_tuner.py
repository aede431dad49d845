"""The served loop's tick tuner (ray-tracer_amd/host/tick_tuner.hpp), driven
on the CPU by a small C++ program against a simulated device whose tick
time is a fixed overhead plus a per-path cost that differs by scene."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DRIVER = r"""
#include <cstdio>
#include "tick_tuner.hpp"
using rtm::TickTuner;
int main() {
  const char* cheap = "c";  // scene keys
  const char* dear = "d";
  // Fixed budget: finals share tick_paths, previews are not charged.
  TickTuner f(1000, 12.0);
  f.begin();
  f.charge_preview(cheap, 5000);
  std::printf("fixed %llu", (unsigned long long)f.final_paths(dear));
  f.charge_final(dear, 600);
  std::printf(" %llu", (unsigned long long)f.final_paths(cheap));
  f.charge_final(cheap, 400);
  std::printf(" %llu\n", (unsigned long long)f.final_paths(cheap));
  // Tuned: the simulated tick costs 0.5 ms + 2e-6 ms/path (cheap) or
  // 2e-5 ms/path (dear), with rays proportional to the cost.
  TickTuner t(0, 12.0, 1u << 20);
  double ms = 0;
  for (int tick = 0; tick < 20; ++tick) {
    t.begin();
    t.charge_preview(cheap, 20000);
    const unsigned long long p = t.final_paths(dear);
    t.charge_final(dear, p);
    ms = 0.5 + 20000 * 2e-6 + p * 2e-5;
    t.add_job(cheap, 20000, 20000 * 2);
    t.add_job(dear, p, p * 20);
    t.end(ms);
  }
  std::printf("tuned %.2f\n", ms);
  // A burst of previews over the target: finals still get a quarter.
  t.begin();
  t.charge_preview(cheap, 100000000);
  std::printf("floor %llu\n", (unsigned long long)t.final_paths(dear));
  t.forget(dear);
  t.begin();
  std::printf("fallback %llu\n", (unsigned long long)t.final_paths(dear));
  return 0;
}
"""


@pytest.mark.skipif(shutil.which("g++") is None, reason="needs g++")
def test_tick_tuner(tmp_path):
    src = tmp_path / "drive.cpp"
    src.write_text(DRIVER)
    exe = tmp_path / "drive"
    subprocess.run(["g++", "-std=c++20", "-O1", "-I", os.path.join(ROOT, "ray-tracer_amd", "host"), str(src), "-o",
                    str(exe)], check=True)
    out = dict(line.split(" ", 1) for line in subprocess.run([str(exe)], capture_output=True, text=True,
                                                               check=True).stdout.splitlines())
    assert out["fixed"].split() == ["1000", "400", "0"]
    # converges to the 12 ms target (the fixed overhead is attributed by rays)
    assert abs(float(out["tuned"]) - 12.0) < 0.6, out["tuned"]
    # a quarter of 12 ms at the dear scene's measured ~2e-5 ms per path
    floor = int(out["floor"])
    assert 100000 < floor < 200000, floor
    # a forgotten scene starts again from the fallback budget (1 Mi paths)
    assert int(out["fallback"]) == (1 << 20)
