// gs_oracle.cpp -- TEST INFRASTRUCTURE ONLY.
//
// CPU restatement of the differentiable Gaussian rasteriser semantics that
// CF-3DGS calls through `diff_gaussian_rasterization` and `simple_knn`
// (imported at scene/gaussian_model_cf.py:20,28-31, called at :873-882 and
// :194-196).  The native sources of those two packages are un-vendored
// submodules (.gitmodules:1-6) and are absent from the reference tree, so
// this file restates the published 3D Gaussian Splatting algorithm (Kerbl et
// al., SIGGRAPH 2023; EWA splatting, Zwicker et al. 2001) plus the Python
// specifications the reference does hold:
//   * SH basis and constants        utils/sh_utils.py:26-112
//   * 3-D covariance from S and q   utils/general_utils.py:76-108,
//                                   scene/gaussian_model_cf.py:42-46
//   * camera matrices               scene/cameras.py:57-74,
//                                   utils/graphics_utils.py:57-113
//   * rigid pose on the means       scene/gaussian_model_cf.py:123-133
//   * kNN initial scale             scene/gaussian_model_cf.py:194-196
//
// It is the checker for the HIP library: only tests/, the smoke entry and the
// CPU-baseline leg of the benchmark load it.  Nothing in the product path may
// call it.  Built twice from one template: fp32 (bit-exact integer spec and
// the image/depth parity target) and fp64 (finite-difference gradient checks).
// Compile with -ffp-contract=off: every rounding step is explicit.
//
// OpenMP parallelism (over Gaussians for per-Gaussian stages, over tiles for
// blending, thread-private gradient buffers in the backward) makes the same
// code the timed CPU baseline.

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <limits>
#include <vector>
#include <omp.h>

namespace {

constexpr int BX = 16;
constexpr int BY = 16;

// ---------------------------------------------------------------------------
// Exponential used by the blend (alpha = o * exp(power)).
//
// Cody & Waite argument reduction (W. J. Cody, W. Waite, "Software Manual for
// the Elementary Functions", 1980) with a degree-7 Taylor polynomial in
// Horner form, every step an explicit fused multiply-add.  fmaf, rintf and
// ldexpf are correctly rounded/exact on both the host and gfx950, so the HIP
// blend computes bit-identical alphas and the 1/255 and 1e-4 decisions agree.
// Accurate to a few ulp over the blend's domain x in [-80, 0].
inline float spec_expf(float x) {
  x = std::fmax(x, -80.0f);
  const float k = std::rint(x * 1.44269502f);
  float r = std::fma(k, -0.693115234375f, x);       // ln2 high part (12 zero tail bits)
  r = std::fma(k, -3.19461833e-05f, r);              // ln2 low part
  float p = 1.98412701e-04f;                         // 1/7!
  p = std::fma(p, r, 1.38888892e-03f);               // 1/6!
  p = std::fma(p, r, 8.33333377e-03f);               // 1/5!
  p = std::fma(p, r, 4.16666679e-02f);               // 1/4!
  p = std::fma(p, r, 1.66666672e-01f);               // 1/3!
  p = std::fma(p, r, 0.5f);
  p = std::fma(p, r, 1.0f);
  p = std::fma(p, r, 1.0f);
  return std::ldexp(p, (int)k);
}

template <typename T> inline T spec_exp(T x);
template <> inline float spec_exp<float>(float x) { return spec_expf(x); }
template <> inline double spec_exp<double>(double x) { return std::exp(x); }

// Saturating float->int conversion (NaN -> 0), the semantics of the GPU's
// convert instruction; keeps the host and device rect arithmetic identical
// for any finite input.
template <typename T> inline int sat_i(T f) {
  if (f != f) return 0;
  if (f >= (T)2147483647.0) return 2147483647;
  if (f <= (T)-2147483648.0) return (-2147483647 - 1);
  return (int)f;
}

// Pixel coordinate from NDC, evaluated in double precision as the upstream
// expression ((v + 1.0) * S - 1.0) * 0.5 promotes to double.
template <typename T> inline T ndc2pix(T v, int S) {
  return (T)((((double)v + 1.0) * (double)S - 1.0) * 0.5);
}

template <typename T> struct V3 { T x, y, z; };

// p' = M p for a column-major 4x4 (the torch row-major transposed matrices of
// scene/cameras.py:58-73).
template <typename T> inline V3<T> xf43(const T* m, const V3<T>& p) {
  return {m[0] * p.x + m[4] * p.y + m[8] * p.z + m[12],
          m[1] * p.x + m[5] * p.y + m[9] * p.z + m[13],
          m[2] * p.x + m[6] * p.y + m[10] * p.z + m[14]};
}
template <typename T> inline T xf44w(const T* m, const V3<T>& p) {
  return m[3] * p.x + m[7] * p.y + m[11] * p.z + m[15];
}

// Rotation matrix of quaternion (r, x, y, z), row-major, exactly the
// polynomial of utils/general_utils.py:88-96 (no normalisation here).
template <typename T> inline void quat_to_R(T r, T x, T y, T z, T R[9]) {
  R[0] = (T)1 - (T)2 * (y * y + z * z);
  R[1] = (T)2 * (x * y - r * z);
  R[2] = (T)2 * (x * z + r * y);
  R[3] = (T)2 * (x * y + r * z);
  R[4] = (T)1 - (T)2 * (x * x + z * z);
  R[5] = (T)2 * (y * z - r * x);
  R[6] = (T)2 * (x * z - r * y);
  R[7] = (T)2 * (y * z + r * x);
  R[8] = (T)1 - (T)2 * (x * x + y * y);
}

// d(loss)/d(r,x,y,z) of the polynomial above, given G = d(loss)/dR (row-major).
template <typename T> inline void quat_R_grad(T r, T x, T y, T z, const T G[9], T dq[4]) {
  dq[0] = (T)2 * (-z * G[1] + y * G[2] + z * G[3] - x * G[5] - y * G[6] + x * G[7]);
  dq[1] = (T)2 * (y * G[1] + z * G[2] + y * G[3] - (T)2 * x * G[4] - r * G[5] + z * G[6] + r * G[7] -
                  (T)2 * x * G[8]);
  dq[2] = (T)2 * (-(T)2 * y * G[0] + x * G[1] + r * G[2] + x * G[3] + z * G[5] - r * G[6] + z * G[7] -
                  (T)2 * y * G[8]);
  dq[3] = (T)2 * (-(T)2 * z * G[0] - r * G[1] + x * G[2] + r * G[3] - (T)2 * z * G[4] + y * G[5] + x * G[6] +
                  y * G[7]);
}

// Pose of layout [tx, ty, tz, qx, qy, qz, qw] (the lietorch SE3 data layout
// used at scene/gaussian_model_cf.py:330-337).  The quaternion is normalised
// before use; p' = R p + t.
template <typename T> struct Pose {
  T R[9];
  T t[3];
  T q[4];  // normalised (w, x, y, z)
  T inv_norm;
};
template <typename T> inline Pose<T> make_pose(const T* p7) {
  Pose<T> P;
  const T qx = p7[3], qy = p7[4], qz = p7[5], qw = p7[6];
  const T n2 = ((qx * qx + qy * qy) + qz * qz) + qw * qw;
  P.inv_norm = (T)1 / std::sqrt(n2);
  P.q[0] = qw * P.inv_norm;
  P.q[1] = qx * P.inv_norm;
  P.q[2] = qy * P.inv_norm;
  P.q[3] = qz * P.inv_norm;
  quat_to_R(P.q[0], P.q[1], P.q[2], P.q[3], P.R);
  P.t[0] = p7[0];
  P.t[1] = p7[1];
  P.t[2] = p7[2];
  return P;
}
template <typename T> inline V3<T> pose_apply(const Pose<T>& P, const V3<T>& p) {
  return {((P.R[0] * p.x + P.R[1] * p.y) + P.R[2] * p.z) + P.t[0],
          ((P.R[3] * p.x + P.R[4] * p.y) + P.R[5] * p.z) + P.t[1],
          ((P.R[6] * p.x + P.R[7] * p.y) + P.R[8] * p.z) + P.t[2]};
}

// SH constants, utils/sh_utils.py:26-45.
template <typename T> struct SHC {
  static constexpr T C0 = (T)0.28209479177387814;
  static constexpr T C1 = (T)0.4886025119029199;
  static constexpr T C2[5] = {(T)1.0925484305920792, (T)-1.0925484305920792, (T)0.31539156525252005,
                              (T)-1.0925484305920792, (T)0.5462742152960396};
  static constexpr T C3[7] = {(T)-0.5900435899266435, (T)2.890611442640554, (T)-0.4570457994644658,
                              (T)0.3731763325901154,  (T)-0.4570457994644658, (T)1.445305721320277,
                              (T)-0.5900435899266435};
};

// The 16 real SH basis values at unit direction (x,y,z); the polynomial of
// utils/sh_utils.py:74-100 with each scalar factor formed before the
// coefficient multiply.
template <typename T> inline void sh_basis(int deg, T x, T y, T z, T b[16]) {
  using C = SHC<T>;
  b[0] = C::C0;
  if (deg < 1) return;
  b[1] = -C::C1 * y;
  b[2] = C::C1 * z;
  b[3] = -C::C1 * x;
  if (deg < 2) return;
  const T xx = x * x, yy = y * y, zz = z * z, xy = x * y, yz = y * z, xz = x * z;
  b[4] = C::C2[0] * xy;
  b[5] = C::C2[1] * yz;
  b[6] = C::C2[2] * (((T)2 * zz - xx) - yy);
  b[7] = C::C2[3] * xz;
  b[8] = C::C2[4] * (xx - yy);
  if (deg < 3) return;
  b[9] = (C::C3[0] * y) * ((T)3 * xx - yy);
  b[10] = (C::C3[1] * xy) * z;
  b[11] = (C::C3[2] * y) * (((T)4 * zz - xx) - yy);
  b[12] = (C::C3[3] * z) * (((T)2 * zz - (T)3 * xx) - (T)3 * yy);
  b[13] = (C::C3[4] * x) * (((T)4 * zz - xx) - yy);
  b[14] = (C::C3[5] * z) * (xx - yy);
  b[15] = (C::C3[6] * x) * (xx - (T)3 * yy);
}

// Gradient of each basis function w.r.t. (x,y,z) of the unit direction.
template <typename T> inline void sh_basis_grad(int deg, T x, T y, T z, T g[16][3]) {
  using C = SHC<T>;
  for (int k = 0; k < 16; ++k) g[k][0] = g[k][1] = g[k][2] = 0;
  if (deg < 1) return;
  g[1][1] = -C::C1;
  g[2][2] = C::C1;
  g[3][0] = -C::C1;
  if (deg < 2) return;
  const T xx = x * x, yy = y * y, zz = z * z, xy = x * y, yz = y * z, xz = x * z;
  g[4][0] = C::C2[0] * y;  g[4][1] = C::C2[0] * x;
  g[5][1] = C::C2[1] * z;  g[5][2] = C::C2[1] * y;
  g[6][0] = -(T)2 * C::C2[2] * x;  g[6][1] = -(T)2 * C::C2[2] * y;  g[6][2] = (T)4 * C::C2[2] * z;
  g[7][0] = C::C2[3] * z;  g[7][2] = C::C2[3] * x;
  g[8][0] = (T)2 * C::C2[4] * x;  g[8][1] = -(T)2 * C::C2[4] * y;
  if (deg < 3) return;
  g[9][0] = C::C3[0] * (T)6 * xy;  g[9][1] = C::C3[0] * ((T)3 * xx - (T)3 * yy);
  g[10][0] = C::C3[1] * yz;  g[10][1] = C::C3[1] * xz;  g[10][2] = C::C3[1] * xy;
  g[11][0] = C::C3[2] * (-(T)2 * xy);  g[11][1] = C::C3[2] * ((T)4 * zz - xx - (T)3 * yy);
  g[11][2] = C::C3[2] * ((T)8 * yz);
  g[12][0] = C::C3[3] * (-(T)6 * xz);  g[12][1] = C::C3[3] * (-(T)6 * yz);
  g[12][2] = C::C3[3] * ((T)6 * zz - (T)3 * xx - (T)3 * yy);
  g[13][0] = C::C3[4] * ((T)4 * zz - (T)3 * xx - yy);  g[13][1] = C::C3[4] * (-(T)2 * xy);
  g[13][2] = C::C3[4] * ((T)8 * xz);
  g[14][0] = C::C3[5] * ((T)2 * xz);  g[14][1] = C::C3[5] * (-(T)2 * yz);  g[14][2] = C::C3[5] * (xx - yy);
  g[15][0] = C::C3[6] * ((T)3 * xx - (T)3 * yy);  g[15][1] = C::C3[6] * (-(T)6 * xy);
}

struct Cam {
  int W, H, D, M, prefiltered;
  double tanfovx, tanfovy, scale_mod;
};

// ---------------------------------------------------------------------------
// Preprocess (per Gaussian).
template <typename T>
void preprocess(int P, const Cam& cam, const T* means, const T* shs, const T* colors, const T* opac,
                const T* scales, const T* rots, const T* cov_pre, const T* view, const T* proj,
                const T* campos, const T* pose7, int* radii, T* xy, T* depth, T* conic_op, T* rgb,
                uint8_t* clamped, uint32_t* tiles, T* cov3D, T* means_tf, int* prefilter_err) {
  const int W = cam.W, H = cam.H;
  const int gx = (W + BX - 1) / BX, gy = (H + BY - 1) / BY;
  const T tanfovx = (T)cam.tanfovx, tanfovy = (T)cam.tanfovy;
  const T fx = (T)W / ((T)2 * tanfovx);
  const T fy = (T)H / ((T)2 * tanfovy);
  const T mod = (T)cam.scale_mod;
  Pose<T> pose{};
  if (pose7) pose = make_pose(pose7);
  int err = 0;
#pragma omp parallel for schedule(static) reduction(| : err)
  for (int i = 0; i < P; ++i) {
    radii[i] = 0;
    tiles[i] = 0;
    V3<T> p = {means[3 * i], means[3 * i + 1], means[3 * i + 2]};
    if (pose7) p = pose_apply(pose, p);
    means_tf[3 * i] = p.x;
    means_tf[3 * i + 1] = p.y;
    means_tf[3 * i + 2] = p.z;
    // near-plane cull on the view-space depth
    const V3<T> pv = xf43(view, p);
    if (pv.z <= (T)0.2) {
      if (cam.prefiltered) err = 1;
      continue;
    }
    const V3<T> ph = xf43(proj, p);
    const T pw = (T)1 / (xf44w(proj, p) + (T)0.0000001);
    const T ndc_x = ph.x * pw, ndc_y = ph.y * pw;

    T* c3 = cov3D + 6 * i;
    if (cov_pre) {
      for (int k = 0; k < 6; ++k) c3[k] = cov_pre[6 * i + k];
    } else {
      T R[9];
      quat_to_R(rots[4 * i], rots[4 * i + 1], rots[4 * i + 2], rots[4 * i + 3], R);
      const T s[3] = {mod * scales[3 * i], mod * scales[3 * i + 1], mod * scales[3 * i + 2]};
      T L[9];
      for (int r = 0; r < 3; ++r)
        for (int k = 0; k < 3; ++k) L[3 * r + k] = R[3 * r + k] * s[k];
      auto sig = [&](int a, int b) { return (L[3 * a] * L[3 * b] + L[3 * a + 1] * L[3 * b + 1]) + L[3 * a + 2] * L[3 * b + 2]; };
      c3[0] = sig(0, 0);
      c3[1] = sig(0, 1);
      c3[2] = sig(0, 2);
      c3[3] = sig(1, 1);
      c3[4] = sig(1, 2);
      c3[5] = sig(2, 2);
    }
    // EWA 2-D covariance: T = J Wr, cov = T Sigma T^T (+0.3 low-pass)
    const T limx = (T)1.3 * tanfovx, limy = (T)1.3 * tanfovy;
    const T txtz = pv.x / pv.z, tytz = pv.y / pv.z;
    const T tx = std::fmin(limx, std::fmax(-limx, txtz)) * pv.z;
    const T ty = std::fmin(limy, std::fmax(-limy, tytz)) * pv.z;
    const T tz = pv.z;
    const T J00 = fx / tz, J02 = -(fx * tx) / (tz * tz);
    const T J11 = fy / tz, J12 = -(fy * ty) / (tz * tz);
    T Tm[2][3];
    for (int c = 0; c < 3; ++c) {
      // Wr[k][c] = view[c*4 + k]
      Tm[0][c] = J00 * view[c * 4 + 0] + J02 * view[c * 4 + 2];
      Tm[1][c] = J11 * view[c * 4 + 1] + J12 * view[c * 4 + 2];
    }
    const T S[3][3] = {{c3[0], c3[1], c3[2]}, {c3[1], c3[3], c3[4]}, {c3[2], c3[4], c3[5]}};
    T MS[2][3];
    for (int a = 0; a < 2; ++a)
      for (int c = 0; c < 3; ++c) MS[a][c] = (Tm[a][0] * S[0][c] + Tm[a][1] * S[1][c]) + Tm[a][2] * S[2][c];
    const T ca = ((MS[0][0] * Tm[0][0] + MS[0][1] * Tm[0][1]) + MS[0][2] * Tm[0][2]) + (T)0.3;
    const T cb = (MS[0][0] * Tm[1][0] + MS[0][1] * Tm[1][1]) + MS[0][2] * Tm[1][2];
    const T cc = ((MS[1][0] * Tm[1][0] + MS[1][1] * Tm[1][1]) + MS[1][2] * Tm[1][2]) + (T)0.3;
    const T det = ca * cc - cb * cb;
    if (det == (T)0) continue;
    const T det_inv = (T)1 / det;
    const T con_a = cc * det_inv, con_b = -cb * det_inv, con_c = ca * det_inv;
    const T mid = (T)0.5 * (ca + cc);
    const T disc = std::sqrt(std::fmax((T)0.1, mid * mid - det));
    const T l1 = mid + disc, l2 = mid - disc;
    const T radius = std::ceil((T)3 * std::sqrt(std::fmax(l1, l2)));
    const T px = ndc2pix(ndc_x, W), py = ndc2pix(ndc_y, H);
    // Non-finite projections are culled (finite inputs never reach this).
    if (!std::isfinite(px) || !std::isfinite(py) || !std::isfinite(radius) || !std::isfinite(con_a) ||
        !std::isfinite(con_b) || !std::isfinite(con_c) || !std::isfinite(pv.z))
      continue;
    const int ir = sat_i(radius);
    const T rr = (T)ir;
    const int x0 = std::min(gx, std::max(0, sat_i((px - rr) / (T)BX)));
    const int y0 = std::min(gy, std::max(0, sat_i((py - rr) / (T)BY)));
    const int x1 = std::min(gx, std::max(0, sat_i((((px + rr) + (T)BX) - (T)1) / (T)BX)));
    const int y1 = std::min(gy, std::max(0, sat_i((((py + rr) + (T)BY) - (T)1) / (T)BY)));
    const uint32_t area = (uint32_t)(x1 - x0) * (uint32_t)(y1 - y0);
    if (area == 0) continue;
    if (shs) {
      const V3<T> dv = {p.x - campos[0], p.y - campos[1], p.z - campos[2]};
      const T len = std::sqrt((dv.x * dv.x + dv.y * dv.y) + dv.z * dv.z);
      const T dx = dv.x / len, dy = dv.y / len, dz = dv.z / len;
      T b[16];
      sh_basis(cam.D, dx, dy, dz, b);
      const int nc = (cam.D + 1) * (cam.D + 1);
      const T* sh = shs + (size_t)i * cam.M * 3;
      for (int ch = 0; ch < 3; ++ch) {
        T acc = b[0] * sh[ch];
        for (int k = 1; k < nc; ++k) acc = acc + b[k] * sh[3 * k + ch];
        acc = acc + (T)0.5;
        clamped[3 * i + ch] = acc < (T)0 ? 1 : 0;
        rgb[3 * i + ch] = std::fmax(acc, (T)0);
      }
    } else {
      for (int ch = 0; ch < 3; ++ch) {
        rgb[3 * i + ch] = colors[3 * i + ch];
        clamped[3 * i + ch] = 0;
      }
    }
    depth[i] = pv.z;
    radii[i] = ir;
    xy[2 * i] = px;
    xy[2 * i + 1] = py;
    conic_op[4 * i] = con_a;
    conic_op[4 * i + 1] = con_b;
    conic_op[4 * i + 2] = con_c;
    conic_op[4 * i + 3] = opac[i];
    tiles[i] = area;
  }
  if (prefilter_err) *prefilter_err = err;
}

inline uint32_t float_bits(float f) {
  uint32_t u;
  std::memcpy(&u, &f, 4);
  return u;
}

// ---------------------------------------------------------------------------
// Binning: duplicate (tile, depth) keys, stable sort, tile ranges.
template <typename T>
long long bin(int P, int W, int H, const int* radii, const T* xy, const T* depth, const uint32_t* tiles,
              uint64_t* keys_dup, uint32_t* vals_dup, uint64_t* keys, uint32_t* vals, uint32_t* ranges) {
  const int gx = (W + BX - 1) / BX, gy = (H + BY - 1) / BY;
  const int Tn = gx * gy;
  std::vector<long long> off(P + 1, 0);
  for (int i = 0; i < P; ++i) off[i + 1] = off[i] + tiles[i];
  const long long n = off[P];
#pragma omp parallel for schedule(static)
  for (int i = 0; i < P; ++i) {
    if (radii[i] <= 0) continue;
    const T px = xy[2 * i], py = xy[2 * i + 1];
    const T rr = (T)radii[i];
    const int x0 = std::min(gx, std::max(0, sat_i((px - rr) / (T)BX)));
    const int y0 = std::min(gy, std::max(0, sat_i((py - rr) / (T)BY)));
    const int x1 = std::min(gx, std::max(0, sat_i((((px + rr) + (T)BX) - (T)1) / (T)BX)));
    const int y1 = std::min(gy, std::max(0, sat_i((((py + rr) + (T)BY) - (T)1) / (T)BY)));
    const uint32_t dbits = float_bits((float)depth[i]);
    long long o = off[i];
    for (int y = y0; y < y1; ++y)
      for (int x = x0; x < x1; ++x) {
        keys_dup[o] = ((uint64_t)(uint32_t)(y * gx + x) << 32) | dbits;
        vals_dup[o] = (uint32_t)i;
        ++o;
      }
  }
  // Stable sort of the 64-bit (tile, depth) keys, in parallel: a stable
  // counting sort on the tile (each thread scatters its own contiguous chunk
  // of the list, in order, from its own per-tile offsets), then a stable
  // sort on the depth bits inside every tile.  The result is the order of a
  // stable sort of the whole keys: equal keys keep their list order.
  const int nthr = std::max(1, omp_get_max_threads());
  std::vector<long long> cnt((size_t)nthr * Tn + 1, 0);
#pragma omp parallel num_threads(nthr)
  {
    const int t = omp_get_thread_num();
    const long long k0 = n * t / nthr, k1 = n * (t + 1) / nthr;
    long long* c = cnt.data() + (size_t)t * Tn;
    for (long long k = k0; k < k1; ++k) ++c[keys_dup[k] >> 32];
  }
  // exclusive prefix in (tile, thread) order
  std::vector<long long> start((size_t)nthr * Tn, 0);
  long long run = 0;
  for (int tile = 0; tile < Tn; ++tile)
    for (int t = 0; t < nthr; ++t) {
      start[(size_t)t * Tn + tile] = run;
      run += cnt[(size_t)t * Tn + tile];
    }
#pragma omp parallel num_threads(nthr)
  {
    const int t = omp_get_thread_num();
    const long long k0 = n * t / nthr, k1 = n * (t + 1) / nthr;
    long long* o = start.data() + (size_t)t * Tn;
    for (long long k = k0; k < k1; ++k) {
      const long long d = o[keys_dup[k] >> 32]++;
      keys[d] = keys_dup[k];
      vals[d] = vals_dup[k];
    }
  }
  std::vector<long long> tile_start(Tn + 1, 0);
  for (int tile = 0; tile < Tn; ++tile) tile_start[tile + 1] = tile_start[tile] + [&] {
      long long c = 0;
      for (int t = 0; t < nthr; ++t) c += cnt[(size_t)t * Tn + tile];
      return c;
    }();
#pragma omp parallel for schedule(dynamic, 1)
  for (int tile = 0; tile < Tn; ++tile) {
    const long long a = tile_start[tile], b = tile_start[tile + 1];
    if (b - a < 2) continue;
    std::vector<std::pair<uint32_t, uint32_t>> kv((size_t)(b - a));
    for (long long k = a; k < b; ++k) kv[k - a] = {(uint32_t)keys[k], vals[k]};
    std::stable_sort(kv.begin(), kv.end(), [](const std::pair<uint32_t, uint32_t>& x,
                                               const std::pair<uint32_t, uint32_t>& y) { return x.first < y.first; });
    for (long long k = a; k < b; ++k) {
      keys[k] = ((uint64_t)(uint32_t)tile << 32) | kv[k - a].first;
      vals[k] = kv[k - a].second;
    }
  }
  std::memset(ranges, 0, sizeof(uint32_t) * 2 * (size_t)Tn);
  for (int tile = 0; tile < Tn; ++tile)
    if (tile_start[tile + 1] > tile_start[tile]) {
      ranges[2 * tile] = (uint32_t)tile_start[tile];
      ranges[2 * tile + 1] = (uint32_t)tile_start[tile + 1];
    }
  return n;
}

// ---------------------------------------------------------------------------
// Per-pixel front-to-back blend.  The alpha, power and accumulation steps
// use the same explicit fused multiply-adds as the HIP kernel.
template <typename T> struct BlendMath;
template <> struct BlendMath<float> {
  static inline float power(float a, float b, float c, float dx, float dy) {
    const float q = std::fma(a * dx, dx, (c * dy) * dy);
    return std::fma(-0.5f, q, -((b * dx) * dy));
  }
};
template <> struct BlendMath<double> {
  static inline double power(double a, double b, double c, double dx, double dy) {
    return -0.5 * (a * dx * dx + c * dy * dy) - b * dx * dy;
  }
};

template <typename T>
void blend(int W, int H, const T* xy, const T* conic_op, const T* rgb, const T* depth, const T* bg,
           const uint32_t* vals, const uint32_t* ranges, T* out_color, T* out_depth, T* out_alpha, T* final_T,
           int* n_contrib) {
  const int gx = (W + BX - 1) / BX, gy = (H + BY - 1) / BY;
  const T a_min = (T)1 / (T)255;
#pragma omp parallel for schedule(dynamic, 1) collapse(2)
  for (int ty = 0; ty < gy; ++ty)
    for (int tx = 0; tx < gx; ++tx) {
      const uint32_t r0 = ranges[2 * (ty * gx + tx)], r1 = ranges[2 * (ty * gx + tx) + 1];
      for (int ly = 0; ly < BY; ++ly)
        for (int lx = 0; lx < BX; ++lx) {
          const int x = tx * BX + lx, y = ty * BY + ly;
          if (x >= W || y >= H) continue;
          const T pfx = (T)x, pfy = (T)y;
          T Tr = 1, C[3] = {0, 0, 0}, Dp = 0;
          int contributor = 0, last = 0;
          for (uint32_t k = r0; k < r1; ++k) {
            ++contributor;
            const uint32_t g = vals[k];
            const T dx = xy[2 * g] - pfx, dy = xy[2 * g + 1] - pfy;
            const T* co = conic_op + 4 * g;
            const T pw = BlendMath<T>::power(co[0], co[1], co[2], dx, dy);
            if (pw > (T)0) continue;
            const T alpha = std::fmin((T)0.99, co[3] * spec_exp<T>(pw));
            if (alpha < a_min) continue;
            const T test_T = Tr * ((T)1 - alpha);
            if (test_T < (T)0.0001) break;
            const T w = alpha * Tr;
            for (int ch = 0; ch < 3; ++ch) C[ch] = std::fma(rgb[3 * g + ch], w, C[ch]);
            Dp = std::fma(depth[g], w, Dp);
            Tr = test_T;
            last = contributor;
          }
          const int pix = y * W + x;
          final_T[pix] = Tr;
          n_contrib[pix] = last;
          for (int ch = 0; ch < 3; ++ch) out_color[ch * H * W + pix] = std::fma(Tr, bg[ch], C[ch]);
          out_depth[pix] = Dp;
          out_alpha[pix] = (T)1 - Tr;
        }
    }
}

// ---------------------------------------------------------------------------
// Backward.
template <typename T>
void backward(int P, const Cam& cam, const T* means, const T* shs, const T* opac, const T* scales,
              const T* rots, const T* cov_pre, const T* view, const T* proj, const T* campos, const T* bg,
              const T* pose7, const int* radii, const T* xy, const T* depth, const T* conic_op, const T* rgb,
              const uint8_t* clamped, const T* cov3D, const T* means_tf, const uint32_t* vals,
              const uint32_t* ranges, const T* final_T, const int* n_contrib, const T* dL_dcolor,
              const T* dL_ddepth, const T* dL_dalpha, T* g_mean2D, T* g_colors, T* g_opac, T* g_means3D,
              T* g_cov3D, T* g_sh, T* g_scales, T* g_rots, T* g_pose7) {
  const int W = cam.W, H = cam.H;
  const int gx = (W + BX - 1) / BX, gy = (H + BY - 1) / BY;
  const T a_min = (T)1 / (T)255;
  const T ddelx = (T)0.5 * (T)W, ddely = (T)0.5 * (T)H;
  // per-Gaussian render gradients: mean2D(2) conic(3: a, b, c) opac color(3) depth
  constexpr int NG = 10;
  const int nthr = omp_get_max_threads();
  std::vector<T> acc((size_t)nthr * P * NG, (T)0);
#pragma omp parallel
  {
    T* my = acc.data() + (size_t)omp_get_thread_num() * P * NG;
#pragma omp for schedule(dynamic, 1) collapse(2)
    for (int ty = 0; ty < gy; ++ty)
      for (int tx = 0; tx < gx; ++tx) {
        const uint32_t r0 = ranges[2 * (ty * gx + tx)], r1 = ranges[2 * (ty * gx + tx) + 1];
        for (int ly = 0; ly < BY; ++ly)
          for (int lx = 0; lx < BX; ++lx) {
            const int x = tx * BX + lx, y = ty * BY + ly;
            if (x >= W || y >= H) continue;
            const int pix = y * W + x;
            const T pfx = (T)x, pfy = (T)y;
            const T T_fin = final_T[pix];
            T Tr = T_fin;
            const int last = n_contrib[pix];
            T dpix[3] = {dL_dcolor[pix], dL_dcolor[H * W + pix], dL_dcolor[2 * H * W + pix]};
            const T dD = dL_ddepth ? dL_ddepth[pix] : (T)0;
            const T dA = dL_dalpha ? dL_dalpha[pix] : (T)0;
            T acc_c[3] = {0, 0, 0}, acc_d = 0, acc_a = 0, last_alpha = 0, last_c[3] = {0, 0, 0}, last_d = 0;
            const T bg_dot = (bg[0] * dpix[0] + bg[1] * dpix[1]) + bg[2] * dpix[2];
            // position k (0-based inside the range) has contributor index k+1
            for (long long k = (long long)r0 + last - 1; k >= (long long)r0; --k) {
              const uint32_t g = vals[k];
              const T dx = xy[2 * g] - pfx, dy = xy[2 * g + 1] - pfy;
              const T* co = conic_op + 4 * g;
              const T pw = BlendMath<T>::power(co[0], co[1], co[2], dx, dy);
              if (pw > (T)0) continue;
              const T G = spec_exp<T>(pw);
              const T alpha = std::fmin((T)0.99, co[3] * G);
              if (alpha < a_min) continue;
              Tr = Tr / ((T)1 - alpha);
              const T w = alpha * Tr;
              T dL_da = 0;
              for (int ch = 0; ch < 3; ++ch) {
                const T c = rgb[3 * g + ch];
                acc_c[ch] = last_alpha * last_c[ch] + ((T)1 - last_alpha) * acc_c[ch];
                last_c[ch] = c;
                dL_da += (c - acc_c[ch]) * dpix[ch];
                my[(size_t)g * NG + 6 + ch] += w * dpix[ch];
              }
              const T z = depth[g];
              acc_d = last_alpha * last_d + ((T)1 - last_alpha) * acc_d;
              last_d = z;
              dL_da += (z - acc_d) * dD;
              my[(size_t)g * NG + 9] += w * dD;
              acc_a = last_alpha + ((T)1 - last_alpha) * acc_a;
              dL_da += ((T)1 - acc_a) * dA;
              dL_da *= Tr;
              last_alpha = alpha;
              dL_da += (-T_fin / ((T)1 - alpha)) * bg_dot;
              const T dL_dG = co[3] * dL_da;
              const T gdx = G * dx, gdy = G * dy;
              const T dG_dx = -gdx * co[0] - gdy * co[1];
              const T dG_dy = -gdy * co[2] - gdx * co[1];
              my[(size_t)g * NG + 0] += dL_dG * dG_dx * ddelx;
              my[(size_t)g * NG + 1] += dL_dG * dG_dy * ddely;
              my[(size_t)g * NG + 2] += (T)-0.5 * gdx * dx * dL_dG;
              my[(size_t)g * NG + 3] += -gdx * dy * dL_dG;
              my[(size_t)g * NG + 4] += (T)-0.5 * gdy * dy * dL_dG;
              my[(size_t)g * NG + 5] += G * dL_da;
            }
          }
      }
  }
  std::vector<T> rg((size_t)P * NG, (T)0);
#pragma omp parallel for schedule(static)
  for (int i = 0; i < P; ++i)
    for (int t = 0; t < nthr; ++t)
      for (int q = 0; q < NG; ++q) rg[(size_t)i * NG + q] += acc[((size_t)t * P + i) * NG + q];

  const T tanfovx = (T)cam.tanfovx, tanfovy = (T)cam.tanfovy;
  const T fx = (T)W / ((T)2 * tanfovx), fy = (T)H / ((T)2 * tanfovy);
  const T mod = (T)cam.scale_mod;
  Pose<T> pose{};
  if (pose7) pose = make_pose(pose7);
  // pose reduction: gt[3], GR[9]
  std::vector<T> pred((size_t)nthr * 12, (T)0);

#pragma omp parallel
  {
    T* mypred = pred.data() + (size_t)omp_get_thread_num() * 12;
#pragma omp for schedule(static)
    for (int i = 0; i < P; ++i) {
      const T* r = rg.data() + (size_t)i * NG;
      g_mean2D[3 * i] = r[0];
      g_mean2D[3 * i + 1] = r[1];
      g_mean2D[3 * i + 2] = 0;
      g_opac[i] = r[5];
      for (int ch = 0; ch < 3; ++ch) g_colors[3 * i + ch] = r[6 + ch];
      for (int k = 0; k < 6; ++k) g_cov3D[6 * i + k] = 0;
      g_means3D[3 * i] = g_means3D[3 * i + 1] = g_means3D[3 * i + 2] = 0;
      if (g_scales) g_scales[3 * i] = g_scales[3 * i + 1] = g_scales[3 * i + 2] = 0;
      if (g_rots) g_rots[4 * i] = g_rots[4 * i + 1] = g_rots[4 * i + 2] = g_rots[4 * i + 3] = 0;
      if (g_sh)
        for (int k = 0; k < cam.M * 3; ++k) g_sh[(size_t)i * cam.M * 3 + k] = 0;
      if (radii[i] <= 0) continue;

      const V3<T> p = {means_tf[3 * i], means_tf[3 * i + 1], means_tf[3 * i + 2]};
      const T* c3 = cov3D + 6 * i;
      // ---- cov2D backward
      const V3<T> t = xf43(view, p);
      const T limx = (T)1.3 * tanfovx, limy = (T)1.3 * tanfovy;
      const T txtz = t.x / t.z, tytz = t.y / t.z;
      const T tx = std::fmin(limx, std::fmax(-limx, txtz)) * t.z;
      const T ty = std::fmin(limy, std::fmax(-limy, tytz)) * t.z;
      const T xm = (txtz < -limx || txtz > limx) ? (T)0 : (T)1;
      const T ym = (tytz < -limy || tytz > limy) ? (T)0 : (T)1;
      const T tz = t.z;
      const T J00 = fx / tz, J02 = -(fx * tx) / (tz * tz);
      const T J11 = fy / tz, J12 = -(fy * ty) / (tz * tz);
      T Tm[2][3];
      for (int c = 0; c < 3; ++c) {
        Tm[0][c] = J00 * view[c * 4 + 0] + J02 * view[c * 4 + 2];
        Tm[1][c] = J11 * view[c * 4 + 1] + J12 * view[c * 4 + 2];
      }
      const T S[3][3] = {{c3[0], c3[1], c3[2]}, {c3[1], c3[3], c3[4]}, {c3[2], c3[4], c3[5]}};
      T ST[2][3];  // (Sigma T_a^T)
      for (int a = 0; a < 2; ++a)
        for (int k = 0; k < 3; ++k) ST[a][k] = S[k][0] * Tm[a][0] + S[k][1] * Tm[a][1] + S[k][2] * Tm[a][2];
      const T ca = Tm[0][0] * ST[0][0] + Tm[0][1] * ST[0][1] + Tm[0][2] * ST[0][2] + (T)0.3;
      const T cb = Tm[0][0] * ST[1][0] + Tm[0][1] * ST[1][1] + Tm[0][2] * ST[1][2];
      const T cc = Tm[1][0] * ST[1][0] + Tm[1][1] * ST[1][1] + Tm[1][2] * ST[1][2] + (T)0.3;
      const T den = ca * cc - cb * cb;
      const T gA = r[2], gB = r[3], gC = r[4];
      const T d2i = (T)1 / (den * den + (T)0.0000001);
      T dLa = 0, dLb = 0, dLc = 0;
      if (d2i != (T)0) {
        dLa = d2i * (-cc * cc * gA + cb * cc * gB + (den - ca * cc) * gC);
        dLc = d2i * (-ca * ca * gC + ca * cb * gB + (den - ca * cc) * gA);
        dLb = d2i * ((T)2 * cb * cc * gA - (den + (T)2 * cb * cb) * gB + (T)2 * ca * cb * gC);
      }
      T gS[6];
      gS[0] = Tm[0][0] * Tm[0][0] * dLa + Tm[0][0] * Tm[1][0] * dLb + Tm[1][0] * Tm[1][0] * dLc;
      gS[3] = Tm[0][1] * Tm[0][1] * dLa + Tm[0][1] * Tm[1][1] * dLb + Tm[1][1] * Tm[1][1] * dLc;
      gS[5] = Tm[0][2] * Tm[0][2] * dLa + Tm[0][2] * Tm[1][2] * dLb + Tm[1][2] * Tm[1][2] * dLc;
      gS[1] = (T)2 * Tm[0][0] * Tm[0][1] * dLa + (Tm[0][0] * Tm[1][1] + Tm[0][1] * Tm[1][0]) * dLb +
              (T)2 * Tm[1][0] * Tm[1][1] * dLc;
      gS[2] = (T)2 * Tm[0][0] * Tm[0][2] * dLa + (Tm[0][0] * Tm[1][2] + Tm[0][2] * Tm[1][0]) * dLb +
              (T)2 * Tm[1][0] * Tm[1][2] * dLc;
      gS[4] = (T)2 * Tm[0][1] * Tm[0][2] * dLa + (Tm[0][1] * Tm[1][2] + Tm[0][2] * Tm[1][1]) * dLb +
              (T)2 * Tm[1][1] * Tm[1][2] * dLc;
      if (d2i == (T)0)
        for (int k = 0; k < 6; ++k) gS[k] = 0;
      // dL/dT
      T gT[2][3];
      for (int k = 0; k < 3; ++k) {
        gT[0][k] = (T)2 * dLa * ST[0][k] + dLb * ST[1][k];
        gT[1][k] = (T)2 * dLc * ST[1][k] + dLb * ST[0][k];
      }
      // dL/dJ (T = J Wr, Wr[k][c] = view[c*4+k])
      T gJ00 = 0, gJ02 = 0, gJ11 = 0, gJ12 = 0;
      for (int c = 0; c < 3; ++c) {
        gJ00 += gT[0][c] * view[c * 4 + 0];
        gJ02 += gT[0][c] * view[c * 4 + 2];
        gJ11 += gT[1][c] * view[c * 4 + 1];
        gJ12 += gT[1][c] * view[c * 4 + 2];
      }
      const T iz = (T)1 / tz, iz2 = iz * iz, iz3 = iz2 * iz;
      const T gtx = xm * -fx * iz2 * gJ02;
      const T gty = ym * -fy * iz2 * gJ12;
      const T gtz = -fx * iz2 * gJ00 - fy * iz2 * gJ11 + ((T)2 * fx * tx) * iz3 * gJ02 + ((T)2 * fy * ty) * iz3 * gJ12;
      // back through t = Wr p + tv
      T gp[3];
      for (int k = 0; k < 3; ++k) gp[k] = view[k * 4 + 0] * gtx + view[k * 4 + 1] * gty + view[k * 4 + 2] * gtz;
      // ---- projection (screen-space mean, NDC units)
      const V3<T> ph = xf43(proj, p);
      const T mw = (T)1 / (xf44w(proj, p) + (T)0.0000001);
      const T mul1 = ph.x * mw * mw, mul2 = ph.y * mw * mw;
      const T g2x = r[0], g2y = r[1];
      for (int k = 0; k < 3; ++k)
        gp[k] += (proj[4 * k + 0] * mw - proj[4 * k + 3] * mul1) * g2x + (proj[4 * k + 1] * mw - proj[4 * k + 3] * mul2) * g2y;
      // ---- depth (= view-space z)
      for (int k = 0; k < 3; ++k) gp[k] += view[4 * k + 2] * r[9];
      // ---- SH colour
      if (shs) {
        const V3<T> dv = {p.x - campos[0], p.y - campos[1], p.z - campos[2]};
        const T len2 = dv.x * dv.x + dv.y * dv.y + dv.z * dv.z;
        const T len = std::sqrt(len2);
        const T dx = dv.x / len, dy = dv.y / len, dz = dv.z / len;
        T gc[3];
        for (int ch = 0; ch < 3; ++ch) gc[ch] = clamped[3 * i + ch] ? (T)0 : r[6 + ch];
        T b[16], bg3[16][3];
        sh_basis(cam.D, dx, dy, dz, b);
        sh_basis_grad(cam.D, dx, dy, dz, bg3);
        const int nc = (cam.D + 1) * (cam.D + 1);
        const T* sh = shs + (size_t)i * cam.M * 3;
        T* gsh = g_sh + (size_t)i * cam.M * 3;
        T gdir[3] = {0, 0, 0};
        for (int k = 0; k < nc; ++k) {
          T s_dot = 0;
          for (int ch = 0; ch < 3; ++ch) {
            gsh[3 * k + ch] = b[k] * gc[ch];
            s_dot += sh[3 * k + ch] * gc[ch];
          }
          for (int a = 0; a < 3; ++a) gdir[a] += s_dot * bg3[k][a];
        }
        // d normalize(v)/dv = (I - n n^T)/|v|
        const T ndot = dx * gdir[0] + dy * gdir[1] + dz * gdir[2];
        gp[0] += (gdir[0] - dx * ndot) / len;
        gp[1] += (gdir[1] - dy * ndot) / len;
        gp[2] += (gdir[2] - dz * ndot) / len;
      }
      // ---- cov3D (scale, rotation) backward
      if (cov_pre) {
        for (int k = 0; k < 6; ++k) g_cov3D[6 * i + k] = gS[k];
      } else {
        for (int k = 0; k < 6; ++k) g_cov3D[6 * i + k] = gS[k];
        const T qr = rots[4 * i], qx = rots[4 * i + 1], qy = rots[4 * i + 2], qz = rots[4 * i + 3];
        T R[9];
        quat_to_R(qr, qx, qy, qz, R);
        const T s[3] = {mod * scales[3 * i], mod * scales[3 * i + 1], mod * scales[3 * i + 2]};
        // G = dL/dSigma (symmetric, off-diagonal halves)
        const T G[9] = {gS[0], (T)0.5 * gS[1], (T)0.5 * gS[2], (T)0.5 * gS[1], gS[3], (T)0.5 * gS[4],
                        (T)0.5 * gS[2], (T)0.5 * gS[4], gS[5]};
        // L = R diag(s); dL/dL = 2 G L
        T L[9], gL[9];
        for (int a = 0; a < 3; ++a)
          for (int k = 0; k < 3; ++k) L[3 * a + k] = R[3 * a + k] * s[k];
        for (int a = 0; a < 3; ++a)
          for (int k = 0; k < 3; ++k)
            gL[3 * a + k] = (T)2 * (G[3 * a] * L[k] + G[3 * a + 1] * L[3 + k] + G[3 * a + 2] * L[6 + k]);
        // dL/ds_k (w.r.t. the modified scale, as the reference does) and dL/dR
        T gR[9];
        for (int k = 0; k < 3; ++k) {
          g_scales[3 * i + k] = gL[k] * R[k] + gL[3 + k] * R[3 + k] + gL[6 + k] * R[6 + k];
          for (int a = 0; a < 3; ++a) gR[3 * a + k] = gL[3 * a + k] * s[k];
        }
        T dq[4];
        quat_R_grad(qr, qx, qy, qz, gR, dq);
        for (int k = 0; k < 4; ++k) g_rots[4 * i + k] = dq[k];
      }
      // ---- pose: p' = R p0 + t
      if (pose7) {
        const T p0[3] = {means[3 * i], means[3 * i + 1], means[3 * i + 2]};
        for (int a = 0; a < 3; ++a) {
          mypred[a] += gp[a];
          for (int k = 0; k < 3; ++k) mypred[3 + 3 * a + k] += gp[a] * p0[k];
        }
        for (int k = 0; k < 3; ++k)
          g_means3D[3 * i + k] = pose.R[k] * gp[0] + pose.R[3 + k] * gp[1] + pose.R[6 + k] * gp[2];
      } else {
        for (int k = 0; k < 3; ++k) g_means3D[3 * i + k] = gp[k];
      }
    }
  }
  if (pose7 && g_pose7) {
    T gt[3] = {0, 0, 0}, GR[9] = {0};
    for (int t = 0; t < nthr; ++t) {
      for (int a = 0; a < 3; ++a) gt[a] += pred[12 * t + a];
      for (int k = 0; k < 9; ++k) GR[k] += pred[12 * t + 3 + k];
    }
    T dqh[4];
    quat_R_grad(pose.q[0], pose.q[1], pose.q[2], pose.q[3], GR, dqh);
    // through the normalisation q_hat = q / |q|
    const T dot = pose.q[0] * dqh[0] + pose.q[1] * dqh[1] + pose.q[2] * dqh[2] + pose.q[3] * dqh[3];
    T dq[4];
    for (int k = 0; k < 4; ++k) dq[k] = (dqh[k] - pose.q[k] * dot) * pose.inv_norm;
    g_pose7[0] = gt[0];
    g_pose7[1] = gt[1];
    g_pose7[2] = gt[2];
    g_pose7[3] = dq[1];
    g_pose7[4] = dq[2];
    g_pose7[5] = dq[3];
    g_pose7[6] = dq[0];
  }
}

// ---------------------------------------------------------------------------
// Exact 3-nearest-neighbour mean squared distance, brute force.
void knn3_brute(int P, const float* pts, float* out) {
#pragma omp parallel for schedule(dynamic, 64)
  for (int i = 0; i < P; ++i) {
    float b0 = std::numeric_limits<float>::max(), b1 = b0, b2 = b0;
    const float x = pts[3 * i], y = pts[3 * i + 1], z = pts[3 * i + 2];
    for (int j = 0; j < P; ++j) {
      if (j == i) continue;
      const float dx = x - pts[3 * j], dy = y - pts[3 * j + 1], dz = z - pts[3 * j + 2];
      const float d = (dx * dx + dy * dy) + dz * dz;
      if (d < b2) {
        if (d < b1) {
          b2 = b1;
          if (d < b0) {
            b1 = b0;
            b0 = d;
          } else {
            b1 = d;
          }
        } else {
          b2 = d;
        }
      }
    }
    out[i] = ((b0 + b1) + b2) / 3.0f;
  }
}

}  // namespace

// ---------------------------------------------------------------------------
// C ABI for the Python test wrapper (oracle/oracle.py).
#define ORC_CAM                                                                                        \
  Cam cam;                                                                                             \
  cam.W = W;                                                                                           \
  cam.H = H;                                                                                           \
  cam.D = D;                                                                                           \
  cam.M = M;                                                                                           \
  cam.prefiltered = prefiltered;                                                                       \
  cam.tanfovx = tanfovx;                                                                               \
  cam.tanfovy = tanfovy;                                                                               \
  cam.scale_mod = scale_mod;

#define ORC_DEFINE(SUF, T)                                                                                   \
  extern "C" int orc_preprocess_##SUF(int P, int D, int M, int W, int H, const T* means, const T* shs,      \
                                      const T* colors, const T* opac, const T* scales, const T* rots,       \
                                      const T* cov_pre, double scale_mod, const T* view, const T* proj,     \
                                      const T* campos, double tanfovx, double tanfovy, const T* pose7,      \
                                      int prefiltered, int* radii, T* xy, T* depth, T* conic_op, T* rgb,    \
                                      uint8_t* clamped, uint32_t* tiles, T* cov3D, T* means_tf) {           \
    ORC_CAM                                                                                                  \
    int err = 0;                                                                                             \
    preprocess<T>(P, cam, means, shs, colors, opac, scales, rots, cov_pre, view, proj, campos, pose7, radii, \
                  xy, depth, conic_op, rgb, clamped, tiles, cov3D, means_tf, &err);                          \
    return err;                                                                                              \
  }                                                                                                          \
  extern "C" long long orc_bin_##SUF(int P, int W, int H, const int* radii, const T* xy, const T* depth,     \
                                     const uint32_t* tiles, uint64_t* keys_dup, uint32_t* vals_dup,          \
                                     uint64_t* keys, uint32_t* vals, uint32_t* ranges) {                     \
    return bin<T>(P, W, H, radii, xy, depth, tiles, keys_dup, vals_dup, keys, vals, ranges);                 \
  }                                                                                                          \
  extern "C" void orc_blend_##SUF(int W, int H, const T* xy, const T* conic_op, const T* rgb,               \
                                  const T* depth, const T* bg, const uint32_t* vals, const uint32_t* ranges, \
                                  T* out_color, T* out_depth, T* out_alpha, T* final_T, int* n_contrib) {    \
    blend<T>(W, H, xy, conic_op, rgb, depth, bg, vals, ranges, out_color, out_depth, out_alpha, final_T,     \
             n_contrib);                                                                                     \
  }                                                                                                          \
  extern "C" void orc_backward_##SUF(                                                                        \
      int P, int D, int M, int W, int H, const T* means, const T* shs, const T* opac, const T* scales,       \
      const T* rots, const T* cov_pre, double scale_mod, const T* view, const T* proj, const T* campos,      \
      const T* bg, double tanfovx, double tanfovy, const T* pose7, const int* radii, const T* xy,            \
      const T* depth, const T* conic_op, const T* rgb, const uint8_t* clamped, const T* cov3D,               \
      const T* means_tf, const uint32_t* vals, const uint32_t* ranges, const T* final_T,                     \
      const int* n_contrib, const T* dL_dcolor, const T* dL_ddepth, const T* dL_dalpha, T* g_mean2D,         \
      T* g_colors, T* g_opac, T* g_means3D, T* g_cov3D, T* g_sh, T* g_scales, T* g_rots, T* g_pose7) {       \
    int prefiltered = 0;                                                                                     \
    ORC_CAM                                                                                                  \
    backward<T>(P, cam, means, shs, opac, scales, rots, cov_pre, view, proj, campos, bg, pose7, radii, xy,   \
                depth, conic_op, rgb, clamped, cov3D, means_tf, vals, ranges, final_T, n_contrib, dL_dcolor, \
                dL_ddepth, dL_dalpha, g_mean2D, g_colors, g_opac, g_means3D, g_cov3D, g_sh, g_scales,        \
                g_rots, g_pose7);                                                                            \
  }

ORC_DEFINE(f32, float)
ORC_DEFINE(f64, double)

extern "C" void orc_knn3_brute(int P, const float* pts, float* out) { knn3_brute(P, pts, out); }

extern "C" void orc_expf(long long n, const float* x, float* y) {
  for (long long i = 0; i < n; ++i) y[i] = spec_expf(x[i]);
}

extern "C" void orc_set_threads(int n) { omp_set_num_threads(n); }
extern "C" int orc_get_threads(void) { return omp_get_max_threads(); }
