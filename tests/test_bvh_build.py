"""The host BVH builder (bvh_build.cpp, replacing Embree's build inside
BuildBVH, src/scene.cpp:209-262) compiled on its own for the host: depth
bound on skewed geometry, every triangle in exactly one leaf. No GPU."""
import os
import subprocess
import tempfile

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CSRC = os.path.join(ROOT, "sh-baker_amd", "csrc")

PROG = r"""
#include "bvh_build.h"
#include <climits>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>
using namespace shb;
int main(int argc, char** argv) {
  const int kind = atoi(argv[1]), width = atoi(argv[3]);
  const uint32_t n = (uint32_t)atoi(argv[2]);
  std::vector<float> xyz(9 * (size_t)n);
  unsigned s = 12345;
  auto rnd = [&]() { s = s * 1103515245u + 12345u; return (float)((s >> 8) & 0xffff) / 65536.f; };
  for (uint32_t i = 0; i < n; ++i) {
    float* t = &xyz[9 * (size_t)i];
    if (kind == 0) {  // sizes and positions halving along x: a chain under the SAH
      const float sc = std::ldexp(1.0f, -(int)(i % 120));
      const float x = sc + 1e-7f * (float)(i / 120);
      const float v[9] = {x, 0, 0, x + sc, 0, 0, x, sc, 0};
      memcpy(t, v, sizeof(v));
    } else {
      for (int k = 0; k < 9; ++k) t[k] = rnd();
    }
  }
  BvhBuildOut out;
  build_cwbvh(xyz.data(), n, width, &out);
  std::vector<int> seen(n, 0);
  for (size_t k = 0; k < out.tris.size(); k += 3) {
    uint32_t id;
    memcpy(&id, &out.tris[k].w, 4);
    if (id >= n) { printf("bad id\n"); return 1; }
    seen[id]++;
  }
  int once = 1;
  for (uint32_t i = 0; i < n; ++i) once &= seen[i] == 1;
  printf("%d %d %d %d\n", out.max_depth, once, out.sah_depth == INT_MAX ? -1 : out.sah_depth,
         kMaxBvhDepth);
  return 0;
}
"""


@pytest.fixture(scope="module")
def builder():
    d = tempfile.mkdtemp()
    src, exe = os.path.join(d, "b.cpp"), os.path.join(d, "b")
    open(src, "w").write(PROG)
    subprocess.run(["g++", "-O2", "-std=c++17", "-DSHB_HOST_ONLY", "-I", CSRC, src,
                    os.path.join(CSRC, "bvh_build.cpp"), "-o", exe, "-lpthread"], check=True)

    def run(kind, n, width):
        out = subprocess.run([exe, str(kind), str(n), str(width)], capture_output=True,
                             text=True, check=True).stdout.split()
        depth, once, budget, limit = map(int, out)
        return depth, bool(once), budget, limit
    return run


@pytest.mark.parametrize("width", [8, 4, 2])
def test_skewed_geometry_stays_within_the_traversal_stack(builder, width):
    depth, once, budget, limit = builder(0, 20000, width)
    assert depth <= limit and once
    assert budget >= 0  # the SAH alone went too deep; a depth budget was used


@pytest.mark.parametrize("width", [8, 4])
def test_ordinary_geometry_keeps_the_sah_tree(builder, width):
    depth, once, budget, limit = builder(1, 30000, width)
    assert depth <= limit and once and budget == -1
