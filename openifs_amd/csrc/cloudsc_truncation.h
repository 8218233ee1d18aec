// Truncation of the explicit sinks of one level (cloudsc.F90:3455-3536) on the CLOUDSC kernel's registers.
//
// The kernel keeps ZSOLQA of one point as its 4 diagonal entries dl, di, dr, ds (ZSOLQA(L,L) .. ZSOLQA(S,S); the
// vapour diagonal is zero on the default path) and 7 antisymmetric pairs, each stored once as x = ZSOLQA(a,b) =
// -ZSOLQA(b,a): il = (I,L), ri = (R,I), rs = (R,S), vl = (V,L), vi = (V,I), vr = (V,R), vs = (V,S); the other
// off-diagonal entries are zero.  Scaling both halves of a pair by the same factor keeps it antisymmetric exactly,
// so one stored value per pair gives the reference's results bit for bit.
//
// Plain C++ apart from the attribute macro, so that the host test (tests/test_truncation.py) compiles it with the
// oracle's truncate_sinks and compares them case by case.
#pragma once

#ifdef __HIPCC__
#define CLOUDSC_HD __host__ __device__
#else
#define CLOUDSC_HD
#endif

namespace cloudsc_gpu {

// qx: ZQX of the level (L, I, R, S, V); div(a, b) and vmax(a, b): the division and MAX of the precision T.
template <typename T, typename Div, typename Max>
CLOUDSC_HD inline void truncate_sinks(T& dl, T& di, T& dr, T& ds, T& il, T& ri, T& rs, T& vl, T& vi, T& vr, T& vs,
                                      const T qx[5], T zepsec, Div div, Max vmax) {
  // Row sums of ZSOLQA (minus sign: sinks positive), in the reference's column order L,I,R,S,V.
  T rsum[5];
  rsum[0] = ((T(0) - dl) + il) + vl;        // row L: (L,L)=dl (L,I)=-il (L,V)=-vl
  rsum[1] = (((T(0) - il) - di) + ri) + vi; // row I: il, di, (I,R)=-ri, (I,V)=-vi
  rsum[2] = (((T(0) - ri) - dr) - rs) + vr; // row R: ri, dr, rs, (R,V)=-vr
  rsum[3] = ((T(0) + rs) - ds) + vs;        // row S: (S,R)=-rs, ds, (S,V)=-vs
  rsum[4] = (((T(0) - vl) - vi) - vr) - vs; // row V: vl, vi, vr, vs
  // ZRATIO = ZMAX/MAX(sinksum,ZMAX) is exactly 1 unless the sinks overshoot; with no
  // overshoot anywhere every scale factor below is 1 and the step is a no-op.
  T zmax[5];
  bool any_trunc = false;
#pragma unroll
  for (int m = 0; m < 5; ++m) {
    zmax[m] = vmax(qx[m], zepsec);
    any_trunc = any_trunc || (rsum[m] > zmax[m]);
  }
  if (any_trunc) {  // lanes without an overshoot are no-ops: every scale factor would be 1
    // ZRATIO: a quotient only for the species that overshoot (liquid and rain nearly always).  `act` marks
    // the species whose step can scale: those that overshoot, and those whose row a scaling has changed
    // since.  Any other species has its first sink sum, at most ZMAX, so its step scales by 1: skipped.
    T zratio[5];
    int act = 0;
#pragma unroll
    for (int m = 0; m < 5; ++m) {
      zratio[m] = T(1);
      if (rsum[m] > zmax[m]) {
        zratio[m] = div(zmax[m], rsum[m]);
        act |= 1 << m;
      }
    }
    // stable ascending order of zratio (insertion sort of :3479-3495): rank[m] = position of species m =
    // #{n < m: zratio[n] <= zratio[m]} + #{n > m: zratio[n] < zratio[m]}, one comparison per pair
    int rank[5] = {0, 0, 0, 0, 0};
#pragma unroll
    for (int m = 1; m < 5; ++m) {
#pragma unroll
      for (int n = 0; n < m; ++n) {
        const bool le = zratio[n] <= zratio[m];
        rank[m] += le ? 1 : 0;
        rank[n] += le ? 0 : 1;
      }
    }
    // Step s processes the species of rank s (:3500-3536): its sink sum from the current entries, in the
    // reference's column order, and on overshoot ZZRATIO = ZMM/ZSINKSUM applied to its sinks ZSOLQA(JO,JN) < 0
    // and their mirror entries ZSOLQA(JN,JO), which marks the rows holding those mirrors.  A stored pair
    // x = ZSOLQA(a,b) = -ZSOLQA(b,a) is a sink of a when x < 0 and of b when x > 0; a negative diagonal entry
    // is scaled twice (:3531-3532).  Each species has its own branch with only its own entries, so a wave
    // executes only the species its lanes process.
#define CLOUDSC_SINK(x, cond) \
if (cond) x = x * zz;
#pragma unroll
    for (int s = 0; s < 5; ++s) {
      if (rank[0] == s && (act & 1)) {  // L: (L,L)=dl, (L,I)=-il, (L,V)=-vl
        const T ssum = ((T(0) - dl) + il) + vl;
        if (ssum > zmax[0]) {
          const T zz = div(zmax[0], ssum);
          CLOUDSC_SINK(il, il > T(0))
          if (dl < T(0)) dl = (dl * zz) * zz;
          CLOUDSC_SINK(vl, vl > T(0))
          act |= 2 | 16;  // rows I (il) and V (vl)
        }
      }
      if (rank[1] == s && (act & 2)) {  // I: (I,L)=il, (I,I)=di, (I,R)=-ri, (I,V)=-vi
        const T ssum = (((T(0) - il) - di) + ri) + vi;
        if (ssum > zmax[1]) {
          const T zz = div(zmax[1], ssum);
          CLOUDSC_SINK(il, il < T(0))
          if (di < T(0)) di = (di * zz) * zz;
          CLOUDSC_SINK(ri, ri > T(0))
          CLOUDSC_SINK(vi, vi > T(0))
          act |= 1 | 4 | 16;  // rows L (il), R (ri), V (vi)
        }
      }
      if (rank[2] == s && (act & 4)) {  // R: (R,I)=ri, (R,R)=dr, (R,S)=rs, (R,V)=-vr
        const T ssum = (((T(0) - ri) - dr) - rs) + vr;
        if (ssum > zmax[2]) {
          const T zz = div(zmax[2], ssum);
          CLOUDSC_SINK(ri, ri < T(0))
          if (dr < T(0)) dr = (dr * zz) * zz;
          CLOUDSC_SINK(rs, rs < T(0))
          CLOUDSC_SINK(vr, vr > T(0))
          act |= 2 | 8 | 16;  // rows I (ri), S (rs), V (vr)
        }
      }
      if (rank[3] == s && (act & 8)) {  // S: (S,R)=-rs, (S,S)=ds, (S,V)=-vs
        const T ssum = ((T(0) + rs) - ds) + vs;
        if (ssum > zmax[3]) {
          const T zz = div(zmax[3], ssum);
          CLOUDSC_SINK(rs, rs > T(0))
          if (ds < T(0)) ds = (ds * zz) * zz;
          CLOUDSC_SINK(vs, vs > T(0))
          act |= 4 | 16;  // rows R (rs), V (vs)
        }
      }
      if (rank[4] == s && (act & 16)) {  // V: (V,L)=vl, (V,I)=vi, (V,R)=vr, (V,S)=vs
        const T ssum = (((T(0) - vl) - vi) - vr) - vs;
        if (ssum > zmax[4]) {
          const T zz = div(zmax[4], ssum);
          CLOUDSC_SINK(vl, vl < T(0))
          CLOUDSC_SINK(vi, vi < T(0))
          CLOUDSC_SINK(vr, vr < T(0))
          CLOUDSC_SINK(vs, vs < T(0))
          act |= 1 | 2 | 4 | 8;  // rows L, I, R, S
        }
      }
    }
#undef CLOUDSC_SINK
  }
}

}  // namespace cloudsc_gpu
