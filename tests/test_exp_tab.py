"""The loss factors' exponential (exp_tab in ram-scb_amd/csrc/sweep_common.h,
Tang's table-driven method) compiled as host code and checked against the C
library's exp over the arguments the loss operators produce and beyond:
within 4 ulp everywhere, 0 below the double range.  The table 2^(j/64) is
the library's own (the same literals the device uses)."""
import re
import shutil
import subprocess
from pathlib import Path

import pytest

REPO = Path(__file__).resolve().parent.parent
CSRC = REPO / "ram-scb_amd" / "csrc"

PROG = r"""
#include <cmath>
#include <cstdio>
#include <random>
#define __host__
#define __device__
#define __forceinline__ inline
%s
int main() {
  const double tab[64] = {%s};
  std::mt19937_64 g(7);
  double worst = 0.0;
  long n = 0;
  auto check = [&](double x) {
    const double a = exp_tab(x, tab), b = std::exp(x);
    if (b < 1e-300) return;  // denormal results: relative error is not meaningful
    const double e = std::fabs(a - b) / b;
    if (e > worst) worst = e;
    ++n;
  };
  std::uniform_real_distribution<double> u(-1.0, 1.0), w(-700.0, 700.0), s(-1e-6, 1e-6);
  for (int i = 0; i < 1000000; ++i) { check(u(g)); check(w(g)); check(s(g)); }
  for (int j = -64; j <= 64; ++j) check(j * 0.6931471805599453 / 64);  // table points
  std::printf("%%ld %%.3e %%.17g %%.17g\n", n, worst, exp_tab(-800.0, tab), exp_tab(0.0, tab));
  return 0;
}
"""


def _source():
    s = (CSRC / "sweep_common.h").read_text()
    i = s.index("__host__ __device__ inline double exp_tab(")
    j = s.index("\n}\n", i) + 3
    fn = s[i:j]
    tab = re.search(r"const double kExp2Tab\[64\] = \{(.*?)\};", (CSRC / "api.hip").read_text(), re.S).group(1)
    return PROG % (fn, tab)


@pytest.mark.skipif(shutil.which("g++") is None, reason="no host C++ compiler")
def test_exp_tab_matches_libm(tmp_path):
    src = tmp_path / "t.cpp"
    src.write_text(_source())
    exe = tmp_path / "t"
    subprocess.run(["g++", "-O2", "-ffp-contract=off", "-std=c++17", str(src), "-o", str(exe)], check=True)
    out = subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout.split()
    n, worst, under, one = int(out[0]), float(out[1]), float(out[2]), float(out[3])
    print(f"{n} arguments, worst relative error {worst:.3e} ({worst / 2.220446049250313e-16:.2f} ulp)")
    assert n > 2_000_000
    assert worst <= 4 * 2.220446049250313e-16
    assert under == 0.0 and one == 1.0
