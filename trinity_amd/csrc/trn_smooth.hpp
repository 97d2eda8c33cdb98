// trn_smooth.hpp -- per-vertex body of the coloured smoothing
// (Smooth::moveSmartLaplacian, src/smoothing.cpp:40-126) of the per-colour
// kernels of smoothing.hip.
//
// The move reads, for the vertex and its 1-ring, coordinates, metric and
// the cached matrix log of the metric, through an accessor over the
// sweeps' working copy (RecVerts: a 64-byte VData per vertex holding the
// log, the coordinates and two metric entries, two vertices per cache
// line, and a dense array of the third metric entry), and the vertex's own
// compact topology (VTopo, 64 bytes).
#pragma once

#include <cfloat>
#include <cstddef>

#include "trn_api.hpp"
#include "trn_numeric.hpp"

namespace trn {

// Per-triangle element quality (Mesh::computeQuality, src/mesh.cpp:626-644)
// from the cached matrix logs of its corners.
__device__ __forceinline__ double quality_logs(const int64_t t, const int32_t* __restrict__ tri,
                                               const double2* __restrict__ xy,
                                               const double4* __restrict__ logm) {
  const int32_t a = tri[3 * t], b = tri[3 * t + 1], c = tri[3 * t + 2];
  double j[3];
  mean_from_logs(logm[a], logm[b], logm[c], j);
  const double2 pa = xy[a], pb = xy[b], pc = xy[c];
  const double A[2] = {pa.x, pa.y}, B[2] = {pb.x, pb.y}, C[2] = {pc.x, pc.y};
  return quality_of_mean(A, B, C, j);
}

// The smoothing sweeps' working copy: vd[j] (log, xy, m11, m12) and t2[j]
// (m22) of every vertex.  Commits also write the new coordinates and metric
// through to the caller's arrays (xo, to) when given, so no copy back is
// needed after the sweeps: those arrays start equal to the working copy
// and change only where a move commits.
struct RecVerts {
  VData* __restrict__ vd;
  double* __restrict__ t2;
  double2* __restrict__ xo;
  double* __restrict__ to;
  __device__ __forceinline__ double2 pos(const int32_t j) const { return vd[j].xy; }
  __device__ __forceinline__ void metric(const int32_t j, double T[3]) const {
    T[0] = vd[j].t01[0];
    T[1] = vd[j].t01[1];
    T[2] = t2[j];
  }
  __device__ __forceinline__ double4 log(const int32_t j) const { return vd[j].logm; }
  __device__ __forceinline__ void commit(const int32_t i, const double px, const double py,
                                         const double TN[3], const double4 LN) const {
    VData d;
    d.logm = LN;
    d.xy = make_double2(px, py);
    d.t01[0] = TN[0];
    d.t01[1] = TN[1];
    vd[i] = d;
    t2[i] = TN[2];
    if (xo) {
      xo[i] = make_double2(px, py);
      to[3 * (int64_t)i] = TN[0];
      to[3 * (int64_t)i + 1] = TN[1];
      to[3 * (int64_t)i + 2] = TN[2];
    }
  }
};

// Topology of a vertex with at most kVtopoDeg incident triangles and 1-ring
// neighbours: ring entries, incident triangles, and per triangle the slot
// of the vertex (rot) and the two other corners (ja, jb).  Every array is
// indexed by unrolled constants only, so the struct lives in registers.
struct FastTopo {
  int deg, nb;
  int32_t ring[kVtopoDeg], tris[kVtopoDeg], ja[kVtopoDeg], jb[kVtopoDeg];
  int rot[kVtopoDeg];
};

// v[k] for a runtime k in [0, 6) as a select chain over named values (a
// runtime-indexed array would be placed in scratch memory).
__device__ __forceinline__ int32_t sel6(const int k, const int32_t v0, const int32_t v1,
                                        const int32_t v2, const int32_t v3, const int32_t v4,
                                        const int32_t v5) {
  int32_t r = v0;
  r = k == 1 ? v1 : r;
  r = k == 2 ? v2 : r;
  r = k == 3 ? v3 : r;
  r = k == 4 ? v4 : r;
  r = k == 5 ? v5 : r;
  return r;
}

// FastTopo from a VTopo record (star bytes r | a << 2 | b << 5).
__device__ __forceinline__ FastTopo fast_topo_rec(const int32_t i, const VTopo& v) {
  FastTopo f;
  f.deg = v.meta & 15;
  f.nb = v.meta >> 4;
#pragma unroll
  for (int k = 0; k < kVtopoDeg; ++k) {
    f.ring[k] = v.ring[k];
    f.tris[k] = v.tris[k];
  }
#pragma unroll
  for (int k = 0; k < kVtopoDeg; ++k) {
    const int s = v.star[k];
    f.rot[k] = s & 3;
    f.ja[k] = sel6((s >> 2) & 7, f.ring[0], f.ring[1], f.ring[2], f.ring[3], f.ring[4],
                   f.ring[5]);
    f.jb[k] = sel6((s >> 5) & 7, f.ring[0], f.ring[1], f.ring[2], f.ring[3], f.ring[4],
                   f.ring[5]);
    if (k >= f.deg) f.ja[k] = f.jb[k] = i;
  }
  return f;
}

// Fast path of the move for a vertex with at most kVtopoDeg incident
// triangles and 1-ring neighbours (every interior vertex of a triangulated
// grid).  The generic loop issues its loads one dependent step at a time
// (CSR entry, then the neighbour, then its data); here all neighbour data
// is loaded in one batch, the averaged tensor of every incident triangle
// is formed once before the first try (it depends on tensors only, so the
// value is the one the lazy form computes, and only triangles a try reaches
// are used), and the corner coordinates stay in registers across the
// tries.  Arithmetic and its order are those of move_vertex_generic.
// lds holds, per thread with stride `stride`, the averaged tensor of each
// incident triangle (3 doubles), its sqrt(det) and the metric length of the
// triangle's edge opposite the vertex: 5 * kVtopoDeg doubles.  Returns 1
// for a commit, 0 for none, 2 when a quality came within 16 ulps of q_min
// (nothing written; the caller redoes the move with the exact quality).
template <class V>
__device__ __forceinline__ int move_vertex_fast(const int32_t i, const FastTopo& ft,
                                                const double2 pini, const double T[3],
                                                const V& v, double* __restrict__ qualit,
                                                double* lds, const int stride) {
  const int deg = ft.deg, nb = ft.nb;
  double q_min = DBL_MAX;  // smoothing.cpp:48-51
#pragma unroll
  for (int k = 0; k < kVtopoDeg; ++k)
    if (k < deg) q_min = dmin(q_min, qualit[ft.tris[k]]);

  const double P[2] = {pini.x, pini.y};
  double opt0 = 0, opt1 = 0;
  double lm[4] = {0., 0., 0., 0.};
#pragma unroll
  for (int k = 0; k < kVtopoDeg; ++k) {  // smoothing.cpp:61-79, numeric.cpp:113-172
    if (k < nb) {
      const int32_t j = ft.ring[k];
      const double2 pb = v.pos(j);
      const double Pb[2] = {pb.x, pb.y};
      double Tb[3];
      v.metric(j, Tb);
      const double4 L = v.log(j);
      const double len = riemann_dist(P, Pb, T, Tb);
      opt0 += P[0] + ((Pb[0] - P[0]) / len);
      opt1 += P[1] + ((Pb[1] - P[1]) / len);
      lm[0] += L.x;
      lm[1] += L.y;
      lm[2] += L.z;
      lm[3] += L.w;
    }
  }
  opt0 /= nb;
  opt1 /= nb;
  double TN[3];
  tensor_exp_mean(lm, nb, TN);  // smoothing.cpp:84
  const double4 LN = tensor_log4(TN[0], TN[1], TN[2]);

  double2 pa[kVtopoDeg], pb[kVtopoDeg];
  double* jm = lds;  // [kVtopoDeg][3], [kVtopoDeg] sqrt(det), [kVtopoDeg] |b - a|
#pragma unroll
  for (int k = 0; k < kVtopoDeg; ++k) {
    if (k < deg) {
      pa[k] = v.pos(ft.ja[k]);
      pb[k] = v.pos(ft.jb[k]);
      const double4 La = v.log(ft.ja[k]), Lb = v.log(ft.jb[k]);
      const int r = ft.rot[k];
      const double4 L0 = r == 0 ? LN : (r == 1 ? Lb : La);
      const double4 L1 = r == 0 ? La : (r == 1 ? LN : Lb);
      const double4 L2 = r == 0 ? Lb : (r == 1 ? La : LN);
      double J[3];
      mean_from_logs(L0, L1, L2, J);
      jm[(3 * k) * stride] = J[0];
      jm[(3 * k + 1) * stride] = J[1];
      jm[(3 * k + 2) * stride] = J[2];
      jm[(3 * kVtopoDeg + k) * stride] = mean_sqrt_det(J);
      jm[(4 * kVtopoDeg + k) * stride] = edge_len(J, pb[k].x - pa[k].x, pb[k].y - pa[k].y);
    } else {
      pa[k] = pb[k] = pini;
    }
  }

  int valid = 0;
  bool undecided = false;
  double w = 1;
  double px = P[0], py = P[1];
  double qs[kVtopoDeg];
  for (int retry = 0; retry < 5; ++retry) {  // smoothing.cpp:96-114
    px = (1 - w) * P[0] + w * opt0;
    py = (1 - w) * P[1] + w * opt1;
    valid = 1;
#pragma unroll
    for (int k = 0; k < kVtopoDeg; ++k) {
      if (valid && k < deg) {
        const int r = ft.rot[k];
        const double2 a = pa[k], b = pb[k];
        // corners in triangle order: v[r] = i, v[r+1] = a, v[r+2] = b
        const double pp[3][2] = {
            {r == 0 ? px : (r == 1 ? b.x : a.x), r == 0 ? py : (r == 1 ? b.y : a.y)},
            {r == 0 ? a.x : (r == 1 ? px : b.x), r == 0 ? a.y : (r == 1 ? py : b.y)},
            {r == 0 ? b.x : (r == 1 ? a.x : px), r == 0 ? b.y : (r == 1 ? a.y : py)}};
        valid = 0;
        if (ccw(pp[0], pp[1], pp[2])) {
          const double J[3] = {jm[(3 * k) * stride], jm[(3 * k + 1) * stride],
                               jm[(3 * k + 2) * stride]};
          // correctly rounded cube: decides as the reference does unless
          // the quality is within 16 ulps of q_min (then the move is left
          // undecided, for the exact path); a stored value may be one ulp
          // from the reference's where glibc's pow(x, 3) is not correctly
          // rounded
          bool near = false;
          qs[k] = quality_with_corner<false>(pp, r, px, py, a, b, J,
                                             jm[(3 * kVtopoDeg + k) * stride],
                                             jm[(4 * kVtopoDeg + k) * stride], q_min, &near);
          undecided |= near;
          valid = !near && qs[k] > q_min;
        }
      }
    }
    w *= 0.5;
    if (valid || undecided) break;
  }
  if (undecided) return 2;
  if (valid) {  // smoothing.cpp:116-123
#pragma unroll
    for (int k = 0; k < kVtopoDeg; ++k)
      if (k < deg) qualit[ft.tris[k]] = qs[k];
    v.commit(i, px, py, TN, LN);
  }
  return valid;
}

// Smooth::moveSmartLaplacian (src/smoothing.cpp:40-126) for vertex i of any
// degree, topology from the CSR arrays.  lds holds, per thread (stride =
// blockDim.x), the averaged tensor (3 doubles) and the quality (1 double) of
// each incident triangle.  The corners of incident triangle k are taken in
// the reference's vertex order (the log-Euclidean mean accumulates in that
// order).
template <bool EXACT, class V>
__device__ __forceinline__ int move_vertex_generic(
    const int32_t i, const int32_t t0, const int deg, const int32_t n0, const int nb,
    const double2 pini, const double T[3], const V& v, double* __restrict__ qualit,
    const int32_t* __restrict__ v2t, const int32_t* __restrict__ star,
    const int32_t* __restrict__ v2v, double* lds, const int stride) {
  const int32_t* ring = v2v + n0;
  double* jm = lds;                      // [deg][3]
  double* qs = lds + 3 * deg * stride;   // [deg]

  double q_min = DBL_MAX;  // smoothing.cpp:48-51
  for (int k = 0; k < deg; ++k) q_min = dmin(q_min, qualit[v2t[t0 + k]]);

  const double P[2] = {pini.x, pini.y};
  double opt0 = 0, opt1 = 0;
  double lm[4] = {0., 0., 0., 0.};
  for (int k = 0; k < nb; ++k) {  // smoothing.cpp:61-79, numeric.cpp:113-172
    const int32_t j = ring[k];
    const double2 pj = v.pos(j);
    const double Pb[2] = {pj.x, pj.y};
    double Tb[3];
    v.metric(j, Tb);
    const double len = riemann_dist(P, Pb, T, Tb);
    opt0 += P[0] + ((Pb[0] - P[0]) / len);
    opt1 += P[1] + ((Pb[1] - P[1]) / len);
    const double4 L = v.log(j);
    lm[0] += L.x;
    lm[1] += L.y;
    lm[2] += L.z;
    lm[3] += L.w;
  }
  opt0 /= nb;
  opt1 /= nb;
  double TN[3];
  tensor_exp_mean(lm, nb, TN);  // smoothing.cpp:84
  const double4 LN = tensor_log4(TN[0], TN[1], TN[2]);

  int valid = 0, nmean = 0;
  double w = 1;
  double px = P[0], py = P[1];
  for (int retry = 0; retry < 5; ++retry) {  // smoothing.cpp:96-114
    px = (1 - w) * P[0] + w * opt0;
    py = (1 - w) * P[1] + w * opt1;
    valid = 1;
    for (int k = 0; valid && k < deg; ++k) {
      const int32_t code = star[t0 + k];
      const int r = code & 3;
      const int32_t ja = ring[(code >> 2) & 255], jb = ring[(code >> 10) & 255];
      const double2 pa = v.pos(ja), pb = v.pos(jb);
      // corners in triangle order: v[r] = i, v[r+1] = a, v[r+2] = b
      // (selects, not a runtime-indexed array, to stay in registers)
      const double pp[3][2] = {
          {r == 0 ? px : (r == 1 ? pb.x : pa.x), r == 0 ? py : (r == 1 ? pb.y : pa.y)},
          {r == 0 ? pa.x : (r == 1 ? px : pb.x), r == 0 ? pa.y : (r == 1 ? py : pb.y)},
          {r == 0 ? pb.x : (r == 1 ? pa.x : px), r == 0 ? pb.y : (r == 1 ? pa.y : py)}};
      valid = 0;
      if (ccw(pp[0], pp[1], pp[2])) {
        double J[3];
        if (k >= nmean) {  // first try to reach triangle k: average once
          const double4 La = v.log(ja), Lb = v.log(jb);
          const double4 L0 = r == 0 ? LN : (r == 1 ? Lb : La);
          const double4 L1 = r == 0 ? La : (r == 1 ? LN : Lb);
          const double4 L2 = r == 0 ? Lb : (r == 1 ? La : LN);
          mean_from_logs(L0, L1, L2, J);
          jm[(3 * k) * stride] = J[0];
          jm[(3 * k + 1) * stride] = J[1];
          jm[(3 * k + 2) * stride] = J[2];
          nmean = k + 1;
        } else {
          J[0] = jm[(3 * k) * stride];
          J[1] = jm[(3 * k + 1) * stride];
          J[2] = jm[(3 * k + 2) * stride];
        }
        // EXACT: the reference's quality; otherwise the correctly rounded
        // cube, and the move is left undecided (2) when a quality comes
        // within 16 ulps of q_min
        bool near = false;
        const double qk = quality_of_mean<EXACT>(pp[0], pp[1], pp[2], J, q_min, &near);
        if (near) return 2;
        qs[k * stride] = qk;
        valid = qk > q_min;
      }
    }
    w *= 0.5;
    if (valid) break;
  }
  if (valid) {  // smoothing.cpp:116-123
    for (int k = 0; k < deg; ++k) qualit[v2t[t0 + k]] = qs[k * stride];
    v.commit(i, px, py, TN, LN);
  }
  return valid;
}

// The move of vertex i with every quality as the reference computes it
// (glibc's pow), through the generic loop; out of line, so that its
// registers do not weigh on the fast path.  Called only for a move the fast
// or generic path left undecided (a quality within 16 ulps of q_min).
static __device__ __noinline__ int move_exact_rec(const int32_t i, const double2 pini,
                                           VData* __restrict__ vd, double* __restrict__ t2,
                                           double2* __restrict__ xo, double* __restrict__ to,
                                           double* __restrict__ qualit,
                                           const int32_t* __restrict__ v2t_off,
                                           const int32_t* __restrict__ v2t,
                                           const int32_t* __restrict__ star,
                                           const int32_t* __restrict__ v2v_off,
                                           const int32_t* __restrict__ v2v, double* lds,
                                           const int stride) {
  const RecVerts v{vd, t2, xo, to};
  const double T[3] = {vd[i].t01[0], vd[i].t01[1], t2[i]};
  const int32_t t0 = v2t_off[i], deg = v2t_off[i + 1] - t0;
  const int32_t n0 = v2v_off[i], nb = v2v_off[i + 1] - n0;
  return move_vertex_generic<true>(i, t0, deg, n0, nb, pini, T, v, qualit, v2t, star, v2v,
                                   lds, stride);
}

// Cooperative form of the fast move for phases with few moves (the live
// lists of the later sweeps, where a phase lasts about one move's latency):
// the G lanes of an aligned group share vertex i, lane s taking the 1-ring
// entries and incident triangles k = s + G j (j < kCoopSlots(G)), so the
// per-neighbour and per-triangle work of the setup and of every try runs
// side by side.  Arithmetic and its order are those of move_vertex_fast:
// sums over the 1-ring are formed in ring order from the lanes' terms
// (every lane adds the same sequence), q_min is the minimum of the cached
// qualities (the reference's fold skips NaN, so NaN enters the group
// minimum as DBL_MAX), and a try's outcome is that of the sequential
// early-exit loop: valid when every triangle passes; otherwise decided by
// the first triangle that fails, undecided (2) when that one came within
// 16 ulps of q_min.  The caller guarantees deg, nb <= kVtopoDeg.  Every
// lane of the group returns the same value; lane 0 commits the vertex data,
// each lane its triangles' qualities.
template <int G>
constexpr int kCoopSlots = (kVtopoDeg + G - 1) / G;

template <int G, class V>
__device__ __forceinline__ int move_vertex_coop(const int32_t i, const int s, const int base,
                                                const int deg, const int nb,
                                                const int32_t (&ring_s)[kCoopSlots<G>],
                                                const int32_t (&tri_s)[kCoopSlots<G>],
                                                const int (&star_s)[kCoopSlots<G>],
                                                const double2 pini, const double T[3],
                                                const V& v, double* __restrict__ qualit) {
  constexpr int S = kCoopSlots<G>;
  const double P[2] = {pini.x, pini.y};
  // q_min (smoothing.cpp:48-51)
  double qm = DBL_MAX;
#pragma unroll
  for (int j = 0; j < S; ++j) {
    if (s + G * j < deg) {
      const double q = qualit[tri_s[j]];
      qm = dmin(qm, q == q ? q : DBL_MAX);
    }
  }
#pragma unroll
  for (int d = 1; d < G; d <<= 1) qm = dmin(qm, __shfl_xor(qm, d, 64));
  const double q_min = qm;
  // 1-ring terms (smoothing.cpp:61-79), summed in ring order
  double ox[S], oy[S];
  double4 L[S];
#pragma unroll
  for (int j = 0; j < S; ++j) {
    ox[j] = oy[j] = 0.;
    L[j] = make_double4(0., 0., 0., 0.);
    if (s + G * j < nb) {
      const double2 pb = v.pos(ring_s[j]);
      const double Pb[2] = {pb.x, pb.y};
      double Tb[3];
      v.metric(ring_s[j], Tb);
      L[j] = v.log(ring_s[j]);
      const double len = riemann_dist(P, Pb, T, Tb);
      ox[j] = P[0] + ((Pb[0] - P[0]) / len);
      oy[j] = P[1] + ((Pb[1] - P[1]) / len);
    }
  }
  double opt0 = 0, opt1 = 0;
  double lm[4] = {0., 0., 0., 0.};
#pragma unroll
  for (int k = 0; k < kVtopoDeg; ++k) {
    const int src = base + k % G, j = k / G;
    const double a0 = __shfl(ox[j], src, 64), a1 = __shfl(oy[j], src, 64);
    const double l0 = __shfl(L[j].x, src, 64), l1 = __shfl(L[j].y, src, 64);
    const double l2 = __shfl(L[j].z, src, 64), l3 = __shfl(L[j].w, src, 64);
    if (k < nb) {
      opt0 += a0;
      opt1 += a1;
      lm[0] += l0;
      lm[1] += l1;
      lm[2] += l2;
      lm[3] += l3;
    }
  }
  opt0 /= nb;
  opt1 /= nb;
  double TN[3];
  tensor_exp_mean(lm, nb, TN);  // smoothing.cpp:84
  const double4 LN = tensor_log4(TN[0], TN[1], TN[2]);
  // the lane's incident triangles: corners, averaged tensor, sqrt(det),
  // opposite edge
  double2 pa[S], pb[S];
  double Jm[S][3], sd[S], el[S], qs[S];
  int rot[S];
#pragma unroll
  for (int j = 0; j < S; ++j) {
    rot[j] = star_s[j] & 3;
    const int ia = (star_s[j] >> 2) & 7, ib = (star_s[j] >> 5) & 7;
    // ring entry q lives in slot q / G of lane q % G
    int32_t ja = i, jb = i;
#pragma unroll
    for (int jj = 0; jj < S; ++jj) {
      const int32_t ca = __shfl(ring_s[jj], base + ia % G, 64);
      const int32_t cb = __shfl(ring_s[jj], base + ib % G, 64);
      if (ia / G == jj) ja = ca;
      if (ib / G == jj) jb = cb;
    }
    pa[j] = pb[j] = pini;
    Jm[j][0] = Jm[j][1] = Jm[j][2] = 0.;
    sd[j] = el[j] = qs[j] = 0.;
    if (s + G * j < deg) {
      pa[j] = v.pos(ja);
      pb[j] = v.pos(jb);
      const double4 La = v.log(ja), Lb = v.log(jb);
      const int r = rot[j];
      const double4 L0 = r == 0 ? LN : (r == 1 ? Lb : La);
      const double4 L1 = r == 0 ? La : (r == 1 ? LN : Lb);
      const double4 L2 = r == 0 ? Lb : (r == 1 ? La : LN);
      mean_from_logs(L0, L1, L2, Jm[j]);
      sd[j] = mean_sqrt_det(Jm[j]);
      el[j] = edge_len(Jm[j], pb[j].x - pa[j].x, pb[j].y - pa[j].y);
    }
  }
  int valid = 0;
  bool undecided = false;
  double w = 1;
  double px = P[0], py = P[1];
  for (int retry = 0; retry < 5; ++retry) {  // smoothing.cpp:96-114
    px = (1 - w) * P[0] + w * opt0;
    py = (1 - w) * P[1] + w * opt1;
    uint32_t badm = 0, nearm = 0;  // bit k: triangle k fails / is near q_min
#pragma unroll
    for (int j = 0; j < S; ++j) {
      const int k = s + G * j;
      if (k < deg) {
        const int r = rot[j];
        const double2 a = pa[j], b = pb[j];
        const double pp[3][2] = {
            {r == 0 ? px : (r == 1 ? b.x : a.x), r == 0 ? py : (r == 1 ? b.y : a.y)},
            {r == 0 ? a.x : (r == 1 ? px : b.x), r == 0 ? a.y : (r == 1 ? py : b.y)},
            {r == 0 ? b.x : (r == 1 ? a.x : px), r == 0 ? b.y : (r == 1 ? a.y : py)}};
        bool ok = false, near = false;
        if (ccw(pp[0], pp[1], pp[2])) {
          qs[j] = quality_with_corner<false>(pp, r, px, py, a, b, Jm[j], sd[j], el[j], q_min,
                                             &near);
          ok = !near && qs[j] > q_min;
        }
        badm |= (uint32_t)!ok << k;
        nearm |= (uint32_t)near << k;
      }
    }
#pragma unroll
    for (int d = 1; d < G; d <<= 1) {
      badm |= (uint32_t)__shfl_xor((int)badm, d, 64);
      nearm |= (uint32_t)__shfl_xor((int)nearm, d, 64);
    }
    valid = badm == 0;
    undecided = !valid && ((nearm >> __builtin_ctz(badm)) & 1);
    w *= 0.5;
    if (valid || undecided) break;
  }
  if (undecided) return 2;
  if (valid) {  // smoothing.cpp:116-123
#pragma unroll
    for (int j = 0; j < S; ++j)
      if (s + G * j < deg) qualit[tri_s[j]] = qs[j];
    if (s == 0) v.commit(i, px, py, TN, LN);
  }
  return valid;
}

// The move on vertex records (per-colour sweeps).  Boundary vertices (tag
// bit 2) and vertices without incident triangles are not moved
// (Smooth::preProcess, src/smoothing.cpp:134-138).  *task = 1 if the
// vertex is a smoothing task; returns 1 if the move was committed.
//
// stale[i] records that the last move of i failed with nothing in its
// closed 1-ring committed since.  A move reads only the closed 1-ring
// (coordinates, metric and logs of i and its neighbours, and the cached
// qualities of its incident triangles, which only a commit at one of their
// corners rewrites), so such a vertex would fail again with the same
// result; the sweeps leave it out of the phase (it still counts as a task).
// A failed move sets stale[i]; a commit clears the flags of the 1-ring.
// In the first sweep (first != 0) every vertex also clears its own flag
// before its move, so no flag is read before it was written in the call.
// Within one phase only the vertex itself and vertices of other colours
// touch stale[i], apart from concurrent commits that all store 0.
__device__ __forceinline__ int move_vertex_rec(const int32_t i, VData* __restrict__ vd,
                                               double* __restrict__ t2,
                                               const VTopo* __restrict__ vtopo,
                                               double2* __restrict__ xo, double* __restrict__ to,
                                               double* __restrict__ qualit,
                                               uint8_t* __restrict__ stale,
                                               const int32_t* __restrict__ v2t_off,
                                               const int32_t* __restrict__ v2t,
                                               const int32_t* __restrict__ star,
                                               const int32_t* __restrict__ v2v_off,
                                               const int32_t* __restrict__ v2v, double* lds,
                                               const int stride, const int first, int* task) {
  const RecVerts v{vd, t2, xo, to};
  if (first) stale[i] = 0;
  const VData own = vd[i];
  const double2 pini = own.xy;
  const double T[3] = {own.t01[0], own.t01[1], t2[i]};
  const VTopo topo = vtopo[i];
  *task = 0;
  if (topo.tag & 4) return 0;
  if (topo.meta != kVtopoGeneric) {
    if ((topo.meta & 15) == 0) return 0;
    *task = 1;
    const FastTopo ft = fast_topo_rec(i, topo);
    int c = move_vertex_fast(i, ft, pini, T, v, qualit, lds, stride);
    if (c == 2)
      c = move_exact_rec(i, pini, vd, t2, xo, to, qualit, v2t_off, v2t, star, v2v_off, v2v,
                         lds, stride);
    if (c) {
#pragma unroll
      for (int k = 0; k < kVtopoDeg; ++k)
        if (k < ft.nb) stale[ft.ring[k]] = 0;
    } else {
      stale[i] = 1;
    }
    return c;
  }
  const int32_t t0 = v2t_off[i], deg = v2t_off[i + 1] - t0;
  const int32_t n0 = v2v_off[i], nb = v2v_off[i + 1] - n0;
  if (deg == 0) return 0;
  *task = 1;
  int c = move_vertex_generic<false>(i, t0, deg, n0, nb, pini, T, v, qualit, v2t, star, v2v,
                                     lds, stride);
  if (c == 2)
    c = move_exact_rec(i, pini, vd, t2, xo, to, qualit, v2t_off, v2t, star, v2v_off, v2v,
                       lds, stride);
  if (c) {
    for (int k = 0; k < nb; ++k) stale[v2v[n0 + k]] = 0;
  } else {
    stale[i] = 1;
  }
  return c;
}

// move_vertex_rec for the G lanes of a group (lane s, group starting at
// lane `base`), every lane holding vertex i: fast-path vertices move
// cooperatively (move_vertex_coop); a move it leaves undecided, and every
// vertex of a larger star, is done by lane 0 alone as move_vertex_rec does
// it.  Every lane returns the same value and task.
template <int G>
__device__ __forceinline__ int move_vertex_rec_coop(
    const int32_t i, const int s, const int base, VData* __restrict__ vd,
    double* __restrict__ t2, const VTopo* __restrict__ vtopo, double2* __restrict__ xo,
    double* __restrict__ to, double* __restrict__ qualit, uint8_t* __restrict__ stale,
    const int32_t* __restrict__ v2t_off, const int32_t* __restrict__ v2t,
    const int32_t* __restrict__ star, const int32_t* __restrict__ v2v_off,
    const int32_t* __restrict__ v2v, double* lds, const int stride, const int first,
    int* task) {
  constexpr int S = kCoopSlots<G>;
  const RecVerts v{vd, t2, xo, to};
  const VTopo* vt = vtopo + i;
  const uint8_t tag = vt->tag, meta = vt->meta;
  *task = 0;
  if (tag & 4) {
    if (first && s == 0) stale[i] = 0;
    return 0;
  }
  int c = 0;
  if (meta != kVtopoGeneric) {
    const int deg = meta & 15, nb = meta >> 4;
    if (first && s == 0) stale[i] = 0;
    if (deg == 0) return 0;
    *task = 1;
    int32_t ring_s[S], tri_s[S];
    int star_s[S];
#pragma unroll
    for (int j = 0; j < S; ++j) {
      const int k = s + G * j;
      ring_s[j] = k < nb ? vt->ring[k] : i;
      tri_s[j] = k < deg ? vt->tris[k] : 0;
      star_s[j] = k < deg ? vt->star[k] : 0;
    }
    const VData own = vd[i];
    const double T[3] = {own.t01[0], own.t01[1], t2[i]};
    c = move_vertex_coop<G>(i, s, base, deg, nb, ring_s, tri_s, star_s, own.xy, T, v, qualit);
    if (c == 2) {
      if (s == 0)
        c = move_exact_rec(i, own.xy, vd, t2, xo, to, qualit, v2t_off, v2t, star, v2v_off, v2v,
                           lds, stride);
      c = __shfl(c, base, 64);
    }
    if (c) {
#pragma unroll
      for (int j = 0; j < S; ++j)
        if (s + G * j < nb) stale[ring_s[j]] = 0;
    } else if (s == 0) {
      stale[i] = 1;
    }
    return c;
  }
  int tk = 0;
  if (s == 0)
    c = move_vertex_rec(i, vd, t2, vtopo, xo, to, qualit, stale, v2t_off, v2t, star, v2v_off,
                        v2v, lds, stride, first, &tk);
  *task = __shfl(tk, base, 64);
  return __shfl(c, base, 64);
}

}  // namespace trn
