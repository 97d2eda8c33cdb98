// trn_smooth.hpp -- per-vertex body of the coloured smoothing
// (Smooth::moveSmartLaplacian, src/smoothing.cpp:40-126), shared by the
// per-colour kernels of smoothing.hip and the persistent batch kernel.
#pragma once

#include <cfloat>

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
  return quality_from_mean(A, B, C, j);
}

// Fast path of move_vertex for a vertex with at most kFastDeg incident
// triangles and 1-ring neighbours (every interior vertex of a triangulated
// grid).  The generic loop issues its loads one dependent step at a time
// (CSR entry, then the neighbour, then its data); here the CSR lines, the
// corner ids of every incident triangle and then all neighbour data are
// each loaded in one batch, the averaged tensor of every incident triangle
// is formed once before the first try (it depends on tensors only, so the
// value is the one the lazy form computes, and only triangles a try reaches
// are used), and the corner coordinates stay in registers across the
// tries.  Arithmetic and its order are those of move_vertex; every array is
// indexed by unrolled constants so it lives in registers.
constexpr int kFastDeg = 6;

__device__ __forceinline__ int move_vertex_fast(
    const int i, const int32_t t0, const int deg, const int32_t n0, const int nb,
    double2* __restrict__ xy, double* __restrict__ tensor, double4* __restrict__ logm,
    double* __restrict__ qualit, const int32_t* __restrict__ v2t,
    const int32_t* __restrict__ star, const int32_t* __restrict__ v2v, double* lds,
    const int stride) {
  int32_t rot[kFastDeg], ja[kFastDeg], jb[kFastDeg];
  double q_min = DBL_MAX;  // smoothing.cpp:48-51
#pragma unroll
  for (int k = 0; k < kFastDeg; ++k) {
    if (k < deg) {
      const int32_t ck = star[t0 + k];
      rot[k] = ck & 3;
      ja[k] = v2v[n0 + ((ck >> 2) & 255)];
      jb[k] = v2v[n0 + ((ck >> 10) & 255)];
      q_min = dmin(q_min, qualit[v2t[t0 + k]]);
    } else {
      rot[k] = 0;
      ja[k] = jb[k] = i;
    }
  }

  const double2 pini = xy[i];
  const double P[2] = {pini.x, pini.y};
  const double T[3] = {tensor[3 * (int64_t)i], tensor[3 * (int64_t)i + 1],
                       tensor[3 * (int64_t)i + 2]};
  double opt0 = 0, opt1 = 0;
  double lm[4] = {0., 0., 0., 0.};
#pragma unroll
  for (int k = 0; k < kFastDeg; ++k) {  // smoothing.cpp:61-79, numeric.cpp:113-172
    if (k < nb) {
      const int32_t j = v2v[n0 + k];
      const double2 pb = xy[j];
      const double Pb[2] = {pb.x, pb.y};
      const double* tb = tensor + 3 * (int64_t)j;
      const double Tb[3] = {tb[0], tb[1], tb[2]};
      const double4 L = logm[j];
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

  double2 pa[kFastDeg], pb[kFastDeg];
  double* jm = lds;  // [kFastDeg][3]
#pragma unroll
  for (int k = 0; k < kFastDeg; ++k) {
    if (k < deg) {
      pa[k] = xy[ja[k]];
      pb[k] = xy[jb[k]];
      const double4 La = logm[ja[k]], Lb = logm[jb[k]];
      const int r = rot[k];
      const double4 L0 = r == 0 ? LN : (r == 1 ? Lb : La);
      const double4 L1 = r == 0 ? La : (r == 1 ? LN : Lb);
      const double4 L2 = r == 0 ? Lb : (r == 1 ? La : LN);
      double J[3];
      mean_from_logs(L0, L1, L2, J);
      jm[(3 * k) * stride] = J[0];
      jm[(3 * k + 1) * stride] = J[1];
      jm[(3 * k + 2) * stride] = J[2];
    } else {
      pa[k] = pb[k] = pini;
    }
  }

  int valid = 0;
  double w = 1;
  double px = P[0], py = P[1];
  double qs[kFastDeg];
  for (int retry = 0; retry < 5; ++retry) {  // smoothing.cpp:96-114
    px = (1 - w) * P[0] + w * opt0;
    py = (1 - w) * P[1] + w * opt1;
    valid = 1;
#pragma unroll
    for (int k = 0; k < kFastDeg; ++k) {
      if (valid && k < deg) {
        const int r = rot[k];
        const double2 a = pa[k], b = pb[k];
        const double pp[3][2] = {
            {r == 0 ? px : (r == 1 ? b.x : a.x), r == 0 ? py : (r == 1 ? b.y : a.y)},
            {r == 0 ? a.x : (r == 1 ? px : b.x), r == 0 ? a.y : (r == 1 ? py : b.y)},
            {r == 0 ? b.x : (r == 1 ? a.x : px), r == 0 ? b.y : (r == 1 ? a.y : py)}};
        valid = 0;
        if (ccw(pp[0], pp[1], pp[2])) {
          const double J[3] = {jm[(3 * k) * stride], jm[(3 * k + 1) * stride],
                               jm[(3 * k + 2) * stride]};
          qs[k] = quality_from_mean(pp[0], pp[1], pp[2], J);
          valid = qs[k] > q_min;
        }
      }
    }
    w *= 0.5;
    if (valid) break;
  }
  if (valid) {  // smoothing.cpp:116-123
#pragma unroll
    for (int k = 0; k < kFastDeg; ++k)
      if (k < deg) qualit[v2t[t0 + k]] = qs[k];
    xy[i] = make_double2(px, py);
    tensor[3 * (int64_t)i] = TN[0];
    tensor[3 * (int64_t)i + 1] = TN[1];
    tensor[3 * (int64_t)i + 2] = TN[2];
    logm[i] = LN;
  }
  return valid;
}

// Smooth::moveSmartLaplacian (src/smoothing.cpp:40-126) for vertex i.
// lds holds, per thread (stride = blockDim.x), the averaged tensor (3
// doubles) and the quality (1 double) of each incident triangle.  The
// corners of incident triangle k are taken in the reference's vertex order
// (the log-Euclidean mean accumulates in that order).
__device__ __forceinline__ int move_vertex(const int i, double2* __restrict__ xy, double* __restrict__ tensor,
                           double4* __restrict__ logm, double* __restrict__ qualit,
                           const int32_t* __restrict__ v2t_off, const int32_t* __restrict__ v2t,
                           const int32_t* __restrict__ star, const int32_t* __restrict__ v2v_off,
                           const int32_t* __restrict__ v2v, double* lds, const int stride) {
  const int32_t t0 = v2t_off[i], deg = v2t_off[i + 1] - t0;
  const int32_t n0 = v2v_off[i], nb = v2v_off[i + 1] - n0;
  if (deg <= kFastDeg && nb <= kFastDeg)
    return move_vertex_fast(i, t0, deg, n0, nb, xy, tensor, logm, qualit, v2t, star, v2v, lds,
                            stride);
  const int32_t* ring = v2v + n0;
  double* jm = lds;                      // [deg][3]
  double* qs = lds + 3 * deg * stride;   // [deg]

  double q_min = DBL_MAX;  // smoothing.cpp:48-51
  for (int k = 0; k < deg; ++k) q_min = dmin(q_min, qualit[v2t[t0 + k]]);

  const double2 pini = xy[i];
  const double P[2] = {pini.x, pini.y};
  const double T[3] = {tensor[3 * (int64_t)i], tensor[3 * (int64_t)i + 1],
                       tensor[3 * (int64_t)i + 2]};
  double opt0 = 0, opt1 = 0;
  double lm[4] = {0., 0., 0., 0.};
  for (int k = 0; k < nb; ++k) {  // smoothing.cpp:61-79, numeric.cpp:113-172
    const int32_t j = ring[k];
    const double2 pj = xy[j];
    const double Pb[2] = {pj.x, pj.y};
    const double* tb = tensor + 3 * (int64_t)j;
    const double Tb[3] = {tb[0], tb[1], tb[2]};
    const double len = riemann_dist(P, Pb, T, Tb);
    opt0 += P[0] + ((Pb[0] - P[0]) / len);
    opt1 += P[1] + ((Pb[1] - P[1]) / len);
    const double4 L = logm[j];
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
      const double2 pa = xy[ja], pb = xy[jb];
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
          const double4 La = logm[ja], Lb = logm[jb];
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
        const double qk = quality_from_mean(pp[0], pp[1], pp[2], J);
        qs[k * stride] = qk;
        valid = qk > q_min;
      }
    }
    w *= 0.5;
    if (valid) break;
  }
  if (valid) {  // smoothing.cpp:116-123
    for (int k = 0; k < deg; ++k) qualit[v2t[t0 + k]] = qs[k * stride];
    xy[i] = make_double2(px, py);
    tensor[3 * (int64_t)i] = TN[0];
    tensor[3 * (int64_t)i + 1] = TN[1];
    tensor[3 * (int64_t)i + 2] = TN[2];
    logm[i] = LN;
  }
  return valid;
}

}  // namespace trn
