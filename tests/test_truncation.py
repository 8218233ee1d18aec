"""The kernel's explicit-sink truncation (openifs_amd/csrc/cloudsc_truncation.h) against the oracle's
(oracle/cloudsc_oracle.hpp truncate_sinks, which follows cloudsc.F90:3455-3536 line by line), bit for bit.

The kernel keeps ZSOLQA as 4 diagonal entries and 7 antisymmetric pairs and skips the steps of species whose row
no scaling has touched.  When no two species below ratio 1 share a pair, their steps commute and it runs them in
species order with their first quotients (the commuting path); otherwise it ranks the species with one comparison
per pair (the sorted path).  Both are compiled here with g++ (the header's device attributes compile away on the
host) and run on seeded cases drawn so that ties between ratios, two species overshooting at once, negative
diagonals, cascades -- a species that overshoots only after another species' scaling has cut one of its sources
-- and both paths all occur many times.  The harness counts those cases itself and the test requires each to be
present.
"""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

HARNESS = r"""
#include <cstdio>
#include <cstdint>
#include <cstring>
#include <random>
#include "oracle/cloudsc_oracle.hpp"
#include "openifs_amd/csrc/cloudsc_truncation.h"

// pair (a,b) -> stored x = ZSOLQA(a,b); species order L I R S V
static const int PA[7][2] = {{1, 0}, {2, 1}, {2, 3}, {4, 0}, {4, 1}, {4, 2}, {4, 3}};

template <typename T>
struct Stats { long cases = 0, bad = 0, trunc = 0, two = 0, ties = 0, cascade = 0, negdiag = 0, partners = 0; };

// the reference loop once more, only to classify the case (which steps scale a species that did not overshoot)
template <typename T>
static bool has_cascade(T qa[5][5], const T qx[5], T zepsec) {
  T zratio[5];
  bool ov[5];
  for (int m = 0; m < 5; ++m) {
    T s = T(0);
    for (int n = 0; n < 5; ++n) s = s - qa[m][n];
    const T zmax = std::max(qx[m], zepsec);
    ov[m] = s > zmax;
    zratio[m] = zmax / std::max(s, zmax);
  }
  int ord[5] = {0, 1, 2, 3, 4};
  std::stable_sort(ord, ord + 5, [&](int a, int b) { return zratio[a] < zratio[b]; });
  bool cascade = false;
  for (int k = 0; k < 5; ++k) {
    const int jo = ord[k];
    T s = T(0);
    for (int n = 0; n < 5; ++n) s = s - qa[jo][n];
    const T zmm = std::max(qx[jo], zepsec);
    const T z = zmm / std::max(s, zmm);
    if (z < T(1) && !ov[jo]) cascade = true;
    for (int n = 0; n < 5; ++n)
      if (qa[jo][n] < T(0)) { qa[jo][n] *= z; qa[n][jo] *= z; }
  }
  return cascade;
}

template <typename T>
static void run(long n, uint64_t seed, Stats<T>& st) {
  std::mt19937_64 g(seed);
  const T mags[6] = {T(0), T(1e-5), T(2e-5), T(3e-6), T(1e-7), T(4e-5)};
  auto val = [&](bool sign) {
    const int k = int(g() % 8);
    T v = k < 6 ? mags[k] : T(std::uniform_real_distribution<double>(0, 5e-5)(g));
    return (sign && (g() & 1)) ? -v : v;
  };
  const T zepsec = T(1e-14);
  for (long c = 0; c < n; ++c) {
    T d[4], x[7], qx[5];
    for (int m = 0; m < 4; ++m) d[m] = (g() % 3 == 0) ? val(true) : T(0);
    for (int p = 0; p < 7; ++p) x[p] = (g() % 4 == 0) ? T(0) : val(true);
    for (int m = 0; m < 5; ++m) qx[m] = (g() % 5 == 0) ? T(0) : val(false);
    T qa[5][5] = {};
    for (int m = 0; m < 4; ++m) qa[m][m] = d[m];
    for (int p = 0; p < 7; ++p) { qa[PA[p][0]][PA[p][1]] = x[p]; qa[PA[p][1]][PA[p][0]] = -x[p]; }
    // classify
    {
      T tmp[5][5];
      std::memcpy(tmp, qa, sizeof tmp);
      T zr[5];
      int nov = 0;
      unsigned first = 0;
      for (int m = 0; m < 5; ++m) {
        T s = T(0);
        for (int k = 0; k < 5; ++k) s = s - qa[m][k];
        const T zmax = std::max(qx[m], zepsec);
        if (s > zmax) ++nov;
        zr[m] = zmax / std::max(s, zmax);
        if (zr[m] < T(1)) first |= 1u << m;
      }
      // species at ratio < 1 that share a pair: the kernel takes its sorted path, otherwise its commuting path
      st.partners += (cloudsc_gpu::truncation_partner_sets() >> first) & 1u;
      bool tie = false;
      for (int a = 0; a < 5; ++a)
        for (int b = a + 1; b < 5; ++b) tie = tie || (zr[a] == zr[b] && zr[a] < T(1));
      st.trunc += nov > 0;
      st.two += nov > 1;
      st.ties += tie;
      st.cascade += has_cascade<T>(tmp, qx, zepsec);
      bool nd = false;
      for (int m = 0; m < 4; ++m) nd = nd || d[m] < T(0);
      st.negdiag += (nov > 0) && nd;
    }
    oracle::truncate_sinks<T>(qa, qx, zepsec);
    T dl = d[0], di = d[1], dr = d[2], ds = d[3];
    T il = x[0], ri = x[1], rs = x[2], vl = x[3], vi = x[4], vr = x[5], vs = x[6];
    cloudsc_gpu::truncate_sinks(dl, di, dr, ds, il, ri, rs, vl, vi, vr, vs, qx, zepsec,
                                [](T a, T b) { return a / b; }, [](T a, T b) { return std::max(a, b); });
    const T got_d[4] = {dl, di, dr, ds}, got_x[7] = {il, ri, rs, vl, vi, vr, vs};
    bool ok = true;
    for (int m = 0; m < 4; ++m) ok = ok && got_d[m] == qa[m][m];
    for (int p = 0; p < 7; ++p)
      ok = ok && got_x[p] == qa[PA[p][0]][PA[p][1]] && -got_x[p] == qa[PA[p][1]][PA[p][0]];
    st.bad += !ok;
    st.cases++;
  }
}

int main() {
  Stats<double> sd;
  Stats<float> sf;
  run<double>(400000, 11, sd);
  run<float>(400000, 12, sf);
  std::printf("double %ld %ld %ld %ld %ld %ld %ld %ld\n", sd.cases, sd.bad, sd.trunc, sd.two, sd.ties, sd.cascade,
              sd.negdiag, sd.partners);
  std::printf("float %ld %ld %ld %ld %ld %ld %ld %ld\n", sf.cases, sf.bad, sf.trunc, sf.two, sf.ties, sf.cascade,
              sf.negdiag, sf.partners);
  return 0;
}
"""


@pytest.fixture(scope="module")
def harness(tmp_path_factory):
    if shutil.which("g++") is None:
        pytest.skip("g++ not available")
    d = tmp_path_factory.mktemp("trunc")
    src = d / "trunc.cpp"
    src.write_text(HARNESS)
    exe = d / "trunc"
    subprocess.run(["g++", "-O1", "-std=c++17", "-ffp-contract=off", "-I", ROOT, str(src), "-o", str(exe)],
                   check=True, timeout=300)
    out = subprocess.run([str(exe)], check=True, capture_output=True, text=True, timeout=300).stdout
    return {line.split()[0]: [int(v) for v in line.split()[1:]] for line in out.strip().splitlines()}


@pytest.mark.parametrize("prec", ["double", "float"])
def test_kernel_truncation_equals_oracle(harness, prec):
    cases, bad, trunc, two, ties, cascade, negdiag, partners = harness[prec]
    print(f"{prec}: {cases} cases, truncating {trunc}, two species {two}, ties {ties}, cascades {cascade}, "
          f"negative diagonal {negdiag}, sorted path {partners}, commuting path {trunc - partners}")
    assert bad == 0, f"{bad} of {cases} cases differ from the oracle"
    # the cases that exercise every branch of the kernel's version are all present
    for name, n in (("truncating", trunc), ("two species", two), ("ties", ties), ("cascades", cascade),
                    ("negative diagonal", negdiag), ("sorted-path", partners), ("commuting-path", trunc - partners)):
        assert n >= 100, f"only {n} {name} cases"
