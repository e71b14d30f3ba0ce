"""The renderer's closed-form empty-space step (advance_steps in
plenvdb_amd/csrc/render.hip) against the loop it replaces,

    while (t + s < tsafe && t < tend && iter < max_iter) { t += s; ++iter; }

in float with round to nearest: the device function's source is compiled for
the host with g++ (-ffp-contract=off, as the library) and run on random
starts, steps, limits and caps -- binade edges, t = 0, steps with few bits
(whose remainders fall exactly half way) and infinite limits included.  The
results (final t and step count) must be equal bit for bit."""
import os
import re
import shutil
import subprocess

import pytest

from conftest import ROOT

HARNESS = r"""
#include <cmath>
#include <cstdio>
#include <random>
#define __device__
#define __forceinline__ inline
using std::fmin; using std::floor; using std::ceil;
%s
int main(int argc, char** argv) {
    std::mt19937_64 g(20261203);
    std::uniform_real_distribution<float> U(0.f, 1.f);
    long bad = 0, n = 0, longer = 0;
    const int cases = atoi(argv[1]);
    for (int it = 0; it < cases; ++it) {
        float t0 = std::ldexp(0.5f + 0.5f * U(g), (int)(g() %% 8) - 4);
        if (g() %% 50 == 0) t0 = (float)(g() %% 4) * 0.25f;
        float s = std::ldexp(1.f + U(g), -(int)(g() %% 16) - 2);
        if (g() %% 20 == 0) s = std::ldexp(1.f + (float)(g() %% 8) / 8.f, -(int)(g() %% 12) - 3);
        float span = std::ldexp(U(g), (int)(g() %% 6) - 2);
        float tsafe = t0 + span, tend = t0 + span * (0.5f + U(g));
        if (g() %% 10 == 0) tend = INFINITY;
        if (g() %% 10 == 0) tsafe = INFINITY;
        int iter0 = (int)(g() %% 100), max_iter = iter0 + (int)(g() %% 200000);
        if ((long)(span / s) > 400000 && tsafe != INFINITY) continue;
        float ta = t0, tb = t0;
        int ia = iter0, ib = iter0;
        advance_steps(ta, s, tsafe, tend, ia, max_iter);
        while (tb + s < tsafe && tb < tend && ib < max_iter) { tb += s; ++ib; }
        ++n;
        if (ib - iter0 > 16) ++longer;
        if (ta != tb || ia != ib) {
            if (bad < 5) printf("mismatch t0=%%a s=%%a tsafe=%%a tend=%%a iter0=%%d max=%%d\n",
                                t0, s, tsafe, tend, iter0, max_iter);
            ++bad;
        }
    }
    printf("%%ld %%ld %%ld\n", n, longer, bad);
    return 0;
}
"""


def _function_source() -> str:
    src = open(os.path.join(ROOT, "plenvdb_amd", "csrc", "render.hip")).read()
    m = re.search(r"__device__ __forceinline__ void advance_steps\(.*?\n}\n", src, re.S)
    assert m, "advance_steps not found in render.hip"
    return m.group(0)


@pytest.mark.skipif(shutil.which("g++") is None, reason="needs g++")
def test_advance_steps_equals_the_step_loop(tmp_path):
    cpp = tmp_path / "adv.cpp"
    exe = tmp_path / "adv"
    cpp.write_text(HARNESS % _function_source())
    subprocess.run(["g++", "-O2", "-ffp-contract=off", "-o", str(exe), str(cpp)], check=True)
    out = subprocess.run([str(exe), "400000"], check=True, capture_output=True, text=True).stdout
    n, longer, bad = (int(x) for x in out.split()[-3:])
    assert n > 300000 and longer > n // 2, out
    assert bad == 0, out
