// CPU restatement (FP32 arithmetic, OpenMP) of the benchmark step: MacCormack
// advection of a 1-channel density with a MAC face velocity (MC_SMOOTH revert
// clamp, BORDER) and the fused density render (BEER_LAMBERT, LINEAR, BORDER)
// with the squared-error image loss, forward and exact adjoint.
//
// TEST INFRASTRUCTURE / CPU BASELINE ONLY.  Nothing in the product (the HIP
// library, the ``ngt`` package) links or calls this file.  It is built by
// oracle/build_cpu_port.py and used by the tests (checked against the FP64
// NumPy oracle) and by the ``cpu_baseline`` leg of bench.py (timed on the
// host's cores).
//
// Semantics follow the NumPy oracle in this directory, which restates:
//  - trilinear reads, BORDER validity 0 <= q + 0.5 < dims, floor / ceil
//    corners, min / max over the 8 corners with the "any invalid corner
//    starts at 0" rule (sampling.hpp:371-704);
//  - VelocityGrid.warp with MacCormack and MC_SMOOTH
//    (data_structures.py:1690-1737) on the centred average of the MAC faces
//    (:1634-1653), sampling at index + ts * v (ts = -dt);
//  - the fused raymarch lattice (raymarch_grid.cu.cc:109-200): start / end at
//    NDC depth 0.5 and zMax + 0.5, BORDER clipping against the cell-centre
//    AABB widened by one cell, steps = floor(length / |step|) + 2, samples
//    read at p - 0.5; BEER_LAMBERT d = max(d + sigma, 0), L += exp(-d) l
//    (blending.hpp:34-39) on the cell (rho * rgb, rho) * depth_step
//    (renderer.py:533-551).
#include <omp.h>

#include <algorithm>
#include <cfloat>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

namespace {

struct Dims3 {
  int x, y, z;  // W, H, D
};

inline int64_t lin(const Dims3& d, int x, int y, int z) { return (static_cast<int64_t>(z) * d.y + y) * d.x + x; }

struct Corners {
  int x0, y0, z0, x1, y1, z1;
  float wx, wy, wz;
  bool vx0, vx1, vy0, vy1, vz0, vz1;
};

inline bool border_inside(const Dims3& d, float qx, float qy, float qz) {
  return qx + 0.5f >= 0.f && qx + 0.5f < static_cast<float>(d.x) && qy + 0.5f >= 0.f &&
         qy + 0.5f < static_cast<float>(d.y) && qz + 0.5f >= 0.f && qz + 0.5f < static_cast<float>(d.z);
}

// only called for positions inside the BORDER range (|q| < dims + 1)
inline Corners corners(const Dims3& d, float qx, float qy, float qz) {
  Corners c;
  const float fx = std::floor(qx), fy = std::floor(qy), fz = std::floor(qz);
  c.x0 = static_cast<int>(fx); c.y0 = static_cast<int>(fy); c.z0 = static_cast<int>(fz);
  c.x1 = static_cast<int>(std::ceil(qx)); c.y1 = static_cast<int>(std::ceil(qy)); c.z1 = static_cast<int>(std::ceil(qz));
  c.wx = qx - fx; c.wy = qy - fy; c.wz = qz - fz;
  c.vx0 = c.x0 > -1 && c.x0 < d.x; c.vx1 = c.x1 > -1 && c.x1 < d.x;
  c.vy0 = c.y0 > -1 && c.y0 < d.y; c.vy1 = c.y1 > -1 && c.y1 < d.y;
  c.vz0 = c.z0 > -1 && c.z0 < d.z; c.vz1 = c.z1 > -1 && c.z1 < d.z;
  return c;
}

// corner k = bx + 2 by + 4 bz (0 = floor, 1 = ceil)
inline bool corner(const Corners& c, int k, int& x, int& y, int& z) {
  const int bx = k & 1, by = (k >> 1) & 1, bz = (k >> 2) & 1;
  x = bx ? c.x1 : c.x0; y = by ? c.y1 : c.y0; z = bz ? c.z1 : c.z0;
  return (bx ? c.vx1 : c.vx0) && (by ? c.vy1 : c.vy0) && (bz ? c.vz1 : c.vz0);
}

inline void fetch8(const float* g, const Dims3& d, const Corners& c, float v[8]) {
  for (int k = 0; k < 8; ++k) {
    int x, y, z;
    v[k] = corner(c, k, x, y, z) ? g[lin(d, x, y, z)] : 0.f;
  }
}

inline float lerp(float a, float b, float t) { return a + t * (b - a); }

float sample_linear(const float* g, const Dims3& d, float qx, float qy, float qz) {
  if (!border_inside(d, qx, qy, qz)) return 0.f;
  const Corners c = corners(d, qx, qy, qz);
  float v[8];
  fetch8(g, d, c, v);
  const float a = lerp(lerp(v[0], v[1], c.wx), lerp(v[2], v[3], c.wx), c.wy);
  const float b = lerp(lerp(v[4], v[5], c.wx), lerp(v[6], v[7], c.wx), c.wy);
  return lerp(a, b, c.wz);
}

void sample_minmax(const float* g, const Dims3& d, float qx, float qy, float qz, float& mn, float& mx) {
  mn = mx = 0.f;
  if (!border_inside(d, qx, qy, qz)) return;
  const Corners c = corners(d, qx, qy, qz);
  const bool all = c.vx0 && c.vx1 && c.vy0 && c.vy1 && c.vz0 && c.vz1;
  mn = all ? FLT_MAX : 0.f;
  mx = all ? -FLT_MAX : 0.f;
  for (int k = 0; k < 8; ++k) {
    int x, y, z;
    if (!corner(c, k, x, y, z)) continue;
    const float v = g[lin(d, x, y, z)];
    mn = std::min(mn, v);
    mx = std::max(mx, v);
  }
}

// d sample_linear / d q (0 outside the BORDER range; 0 along an axis at
// integer q, where floor and ceil coincide)
void data_gradient(const float* g, const Dims3& d, float qx, float qy, float qz, float& gx, float& gy, float& gz) {
  gx = gy = gz = 0.f;
  if (!border_inside(d, qx, qy, qz)) return;
  const Corners c = corners(d, qx, qy, qz);
  float v[8];
  fetch8(g, d, c, v);
  gx = lerp(lerp(v[1] - v[0], v[3] - v[2], c.wy), lerp(v[5] - v[4], v[7] - v[6], c.wy), c.wz);
  gy = lerp(lerp(v[2] - v[0], v[3] - v[1], c.wx), lerp(v[6] - v[4], v[7] - v[5], c.wx), c.wz);
  gz = lerp(lerp(v[4] - v[0], v[5] - v[1], c.wx), lerp(v[6] - v[2], v[7] - v[3], c.wx), c.wy);
}

inline void atomic_add(float* p, float v) {
#pragma omp atomic
  *p += v;
}

void scatter_linear(float* g, const Dims3& d, float qx, float qy, float qz, float val) {
  if (!border_inside(d, qx, qy, qz)) return;
  const Corners c = corners(d, qx, qy, qz);
  const float fw[2][3] = {{1.f - c.wx, 1.f - c.wy, 1.f - c.wz}, {c.wx, c.wy, c.wz}};
  for (int k = 0; k < 8; ++k) {
    int x, y, z;
    if (!corner(c, k, x, y, z)) continue;
    const float w = fw[k & 1][0] * fw[(k >> 1) & 1][1] * fw[(k >> 2) & 1][2];
    atomic_add(g + lin(d, x, y, z), val * w);
  }
}

// centred velocity of cell (x, y, z) from the MAC faces
inline void centered(const float* vx, const float* vy, const float* vz, const Dims3& d, int x, int y, int z, float& cx,
                     float& cy, float& cz) {
  const Dims3 dx{d.x + 1, d.y, d.z}, dy{d.x, d.y + 1, d.z}, dz{d.x, d.y, d.z + 1};
  cx = 0.5f * (vx[lin(dx, x + 1, y, z)] + vx[lin(dx, x, y, z)]);
  cy = 0.5f * (vy[lin(dy, x, y + 1, z)] + vy[lin(dy, x, y, z)]);
  cz = 0.5f * (vz[lin(dz, x, y, z + 1)] + vz[lin(dz, x, y, z)]);
}

// ---- cameras ----------------------------------------------------------------

using Mat = double[4][4];

void load_math(const float* flat, Mat m) {  // op argument (column-major) -> math matrix
  for (int r = 0; r < 4; ++r)
    for (int c = 0; c < 4; ++c) m[r][c] = flat[c * 4 + r];
}

void matmul(const Mat a, const Mat b, Mat o) {
  for (int r = 0; r < 4; ++r)
    for (int c = 0; c < 4; ++c) {
      double s = 0.0;
      for (int k = 0; k < 4; ++k) s += a[r][k] * b[k][c];
      o[r][c] = s;
    }
}

bool invert(const Mat a, Mat o) {  // Gauss-Jordan, partial pivoting
  double m[4][8];
  for (int r = 0; r < 4; ++r)
    for (int c = 0; c < 4; ++c) {
      m[r][c] = a[r][c];
      m[r][c + 4] = r == c ? 1.0 : 0.0;
    }
  for (int col = 0; col < 4; ++col) {
    int piv = col;
    for (int r = col + 1; r < 4; ++r)
      if (std::fabs(m[r][col]) > std::fabs(m[piv][col])) piv = r;
    if (std::fabs(m[piv][col]) < 1e-300) return false;
    if (piv != col)
      for (int c = 0; c < 8; ++c) std::swap(m[piv][c], m[col][c]);
    const double inv = 1.0 / m[col][col];
    for (int c = 0; c < 8; ++c) m[col][c] *= inv;
    for (int r = 0; r < 4; ++r) {
      if (r == col) continue;
      const double f = m[r][col];
      for (int c = 0; c < 8; ++c) m[r][c] -= f * m[col][c];
    }
  }
  for (int r = 0; r < 4; ++r)
    for (int c = 0; c < 4; ++c) o[r][c] = m[r][c + 4];
  return true;
}

struct Camera {
  Mat mv_inv, p_inv;
};

Camera make_camera(const float* m16, const float* v16, const float* p16) {
  Mat M, V, P, MV;
  load_math(m16, M);
  load_math(v16, V);
  load_math(p16, P);
  matmul(V, M, MV);
  Camera c;
  invert(MV, c.mv_inv);
  invert(P, c.p_inv);
  return c;
}

// IDXtoOS: index coordinates of the frustum grid -> object space
void idx_to_os(const Camera& c, double ix, double iy, double iz, int W, int H, int D, double o[3]) {
  const double ndc[4] = {2.0 * ix / W - 1.0, 2.0 * iy / H - 1.0, 2.0 * iz / D - 1.0, 1.0};
  double vs[4];
  for (int r = 0; r < 4; ++r) vs[r] = c.p_inv[r][0] * ndc[0] + c.p_inv[r][1] * ndc[1] + c.p_inv[r][2] * ndc[2] + c.p_inv[r][3];
  for (int r = 0; r < 4; ++r) vs[r] /= vs[3];
  for (int r = 0; r < 3; ++r)
    o[r] = c.mv_inv[r][0] * vs[0] + c.mv_inv[r][1] * vs[1] + c.mv_inv[r][2] * vs[2] + c.mv_inv[r][3] * vs[3];
}

struct Ray {
  float sx, sy, sz, dx, dy, dz;
  int steps;
};

Ray make_ray(const Camera& c, const Dims3& din, int px, int py, int W, int H, int D) {
  const double zmax = std::max(D - 1, 1);
  double s[3], e[3];
  idx_to_os(c, px + 0.5, py + 0.5, 0.5, W, H, D, s);
  idx_to_os(c, px + 0.5, py + 0.5, zmax + 0.5, W, H, D, e);
  double step[3], seg[3];
  for (int a = 0; a < 3; ++a) {
    step[a] = (e[a] - s[a]) / zmax;
    seg[a] = e[a] - s[a];
  }
  const double len = std::sqrt(seg[0] * seg[0] + seg[1] * seg[1] + seg[2] * seg[2]);
  const double dims[3] = {static_cast<double>(din.x), static_cast<double>(din.y), static_cast<double>(din.z)};
  double tmin = -INFINITY, tmax = INFINITY, rdir[3];
  for (int a = 0; a < 3; ++a) {
    rdir[a] = seg[a] / len;
    const double inv = 1.0 / rdir[a];
    const double n = inv * (s[a] - dims[a] * 0.5), k = std::fabs(inv) * (dims[a] * 0.5 - 0.5);
    tmin = std::fmax(tmin, -n - k);  // fmax / fmin skip the NaN of an axis-parallel ray
    tmax = std::fmin(tmax, k - n);
  }
  Ray r{};
  if (tmin > tmax || tmax < 0.0) return r;
  double st[3], en[3];
  for (int a = 0; a < 3; ++a) {
    en[a] = s[a] + rdir[a] * (tmax + 1.0);
    st[a] = s[a] + rdir[a] * std::max(tmin - 1.0, 0.0);
  }
  const double ss = std::sqrt(step[0] * step[0] + step[1] * step[1] + step[2] * step[2]);
  const double l = std::sqrt((st[0] - en[0]) * (st[0] - en[0]) + (st[1] - en[1]) * (st[1] - en[1]) +
                             (st[2] - en[2]) * (st[2] - en[2]));
  r.sx = static_cast<float>(st[0]); r.sy = static_cast<float>(st[1]); r.sz = static_cast<float>(st[2]);
  r.dx = static_cast<float>(step[0]); r.dy = static_cast<float>(step[1]); r.dz = static_cast<float>(step[2]);
  r.steps = static_cast<int>(std::floor(l / ss)) + 2;
  return r;
}

}  // namespace

extern "C" {

int ngtcpu_threads(void) { return omp_get_max_threads(); }

// MacCormack step: sl = S(rho, i + ts v), out = MC_SMOOTH(sl + (rho - S(sl, i - ts v)) / 2).
void ngtcpu_advect_mc_fwd(const float* rho, int D, int H, int W, const float* vx, const float* vy, const float* vz,
                          float ts, float* sl, float* out) {
  const Dims3 d{W, H, D};
#pragma omp parallel for collapse(2) schedule(static)
  for (int z = 0; z < D; ++z)
    for (int y = 0; y < H; ++y)
      for (int x = 0; x < W; ++x) {
        float cx, cy, cz;
        centered(vx, vy, vz, d, x, y, z, cx, cy, cz);
        sl[lin(d, x, y, z)] = sample_linear(rho, d, x + ts * cx, y + ts * cy, z + ts * cz);
      }
#pragma omp parallel for collapse(2) schedule(static)
  for (int z = 0; z < D; ++z)
    for (int y = 0; y < H; ++y)
      for (int x = 0; x < W; ++x) {
        float cx, cy, cz;
        centered(vx, vy, vz, d, x, y, z, cx, cy, cz);
        const int64_t i = lin(d, x, y, z);
        const float back = sample_linear(sl, d, x - ts * cx, y - ts * cy, z - ts * cz);
        const float corr = sl[i] + 0.5f * (rho[i] - back);
        float mn, mx;
        sample_minmax(rho, d, x + ts * cx, y + ts * cy, z + ts * cz, mn, mx);
        out[i] = (corr < mn || corr > mx) ? sl[i] : corr;
      }
}

// Exact adjoint of ngtcpu_advect_mc_fwd (sl = its semi-Lagrangian field).
// work: 4 * D*H*W floats.  g_vx / g_vy / g_vz: MAC face gradients.
void ngtcpu_advect_mc_bwd(const float* rho, int D, int H, int W, const float* vx, const float* vy, const float* vz,
                          float ts, const float* sl, const float* g_out, float* g_rho, float* g_vx, float* g_vy,
                          float* g_vz, float* work) {
  const Dims3 d{W, H, D};
  const int64_t cells = static_cast<int64_t>(D) * H * W;
  float* g_sl = work;
  float* gv = work + cells;  // centred velocity gradient, 3 planes
  std::memcpy(g_sl, g_out, cells * sizeof(float));
  std::memset(gv, 0, 3 * cells * sizeof(float));
#pragma omp parallel for collapse(2) schedule(static)
  for (int z = 0; z < D; ++z)
    for (int y = 0; y < H; ++y)
      for (int x = 0; x < W; ++x) {
        float cx, cy, cz;
        centered(vx, vy, vz, d, x, y, z, cx, cy, cz);
        const int64_t i = lin(d, x, y, z);
        const float bx = x - ts * cx, by = y - ts * cy, bz = z - ts * cz;
        const float back = sample_linear(sl, d, bx, by, bz);
        const float corr = sl[i] + 0.5f * (rho[i] - back);
        float mn, mx;
        sample_minmax(rho, d, x + ts * cx, y + ts * cy, z + ts * cz, mn, mx);
        const float g_corr = (corr < mn || corr > mx) ? 0.f : g_out[i];
        g_rho[i] = 0.5f * g_corr;
        const float g_back = -0.5f * g_corr;
        scatter_linear(g_sl, d, bx, by, bz, g_back);
        float dgx, dgy, dgz;
        data_gradient(sl, d, bx, by, bz, dgx, dgy, dgz);
        // d back / d v = -ts * grad(sl)
        gv[i] -= ts * dgx * g_back;
        gv[cells + i] -= ts * dgy * g_back;
        gv[2 * cells + i] -= ts * dgz * g_back;
      }
#pragma omp parallel for collapse(2) schedule(static)
  for (int z = 0; z < D; ++z)
    for (int y = 0; y < H; ++y)
      for (int x = 0; x < W; ++x) {
        float cx, cy, cz;
        centered(vx, vy, vz, d, x, y, z, cx, cy, cz);
        const int64_t i = lin(d, x, y, z);
        const float px = x + ts * cx, py = y + ts * cy, pz = z + ts * cz;
        scatter_linear(g_rho, d, px, py, pz, g_sl[i]);
        float dgx, dgy, dgz;
        data_gradient(rho, d, px, py, pz, dgx, dgy, dgz);
        gv[i] += ts * dgx * g_sl[i];
        gv[cells + i] += ts * dgy * g_sl[i];
        gv[2 * cells + i] += ts * dgz * g_sl[i];
      }
  // centred average -> faces
  const Dims3 dx{W + 1, H, D}, dy{W, H + 1, D}, dz{W, H, D + 1};
#pragma omp parallel for collapse(2) schedule(static)
  for (int z = 0; z < D + 1; ++z)
    for (int y = 0; y < H + 1; ++y)
      for (int x = 0; x < W + 1; ++x) {
        if (z < D && y < H) {
          float s = 0.f;
          if (x > 0) s += gv[lin(d, x - 1, y, z)];
          if (x < W) s += gv[lin(d, x, y, z)];
          g_vx[lin(dx, x, y, z)] = 0.5f * s;
        }
        if (z < D && x < W) {
          float s = 0.f;
          if (y > 0) s += gv[cells + lin(d, x, y - 1, z)];
          if (y < H) s += gv[cells + lin(d, x, y, z)];
          g_vy[lin(dy, x, y, z)] = 0.5f * s;
        }
        if (y < H && x < W) {
          float s = 0.f;
          if (z > 0) s += gv[2 * cells + lin(d, x, y, z - 1)];
          if (z < D) s += gv[2 * cells + lin(d, x, y, z)];
          g_vz[lin(dz, x, y, z)] = 0.5f * s;
        }
      }
}

// Fused density render: images [V, OH, OW, 4] = (L_rgb, optical depth).
void ngtcpu_render_fwd(const float* rho, int D, int H, int W, const float* m16, const float* v16, const float* p16,
                       int views, int OD, int OH, int OW, const float* light_rgb, float depth_step, float* images) {
  const Dims3 d{W, H, D};
  for (int v = 0; v < views; ++v) {
    const Camera cam = make_camera(m16, v16 + 16 * v, p16 + 16 * v);
#pragma omp parallel for collapse(2) schedule(dynamic, 4)
    for (int py = 0; py < OH; ++py)
      for (int px = 0; px < OW; ++px) {
        const Ray r = make_ray(cam, d, px, py, OW, OH, OD);
        float L[3] = {0.f, 0.f, 0.f}, dep = 0.f;
        for (int k = 0; k < r.steps; ++k) {
          const float kf = static_cast<float>(k);
          const float s = depth_step * sample_linear(rho, d, std::fma(kf, r.dx, r.sx) - 0.5f,
                                                     std::fma(kf, r.dy, r.sy) - 0.5f, std::fma(kf, r.dz, r.sz) - 0.5f);
          dep = std::max(dep + s, 0.f);
          const float e = std::exp(-dep);
          for (int c = 0; c < 3; ++c) L[c] += e * light_rgb[c] * s;
        }
        float* o = images + ((static_cast<int64_t>(v) * OH + py) * OW + px) * 4;
        o[0] = L[0]; o[1] = L[1]; o[2] = L[2]; o[3] = dep;
      }
  }
}

// loss_weight * sum over pixels and RGB of (image - target)^2.
double ngtcpu_l2_loss(const float* images, const float* target, int64_t npix, float loss_weight) {
  double s = 0.0;
#pragma omp parallel for reduction(+ : s) schedule(static)
  for (int64_t p = 0; p < npix; ++p)
    for (int c = 0; c < 3; ++c) {
      const double e = static_cast<double>(images[p * 4 + c]) - target[p * 3 + c];
      s += e * e;
    }
  return loss_weight * s;
}

// Exact adjoint of render + L2: density gradient (overwritten) for the loss
// gradient loss_grad.
void ngtcpu_render_l2_bwd(const float* rho, int D, int H, int W, const float* m16, const float* v16, const float* p16,
                          int views, int OD, int OH, int OW, const float* light_rgb, float depth_step,
                          const float* images, const float* target, float loss_weight, float loss_grad,
                          float* g_rho) {
  const Dims3 d{W, H, D};
  std::memset(g_rho, 0, static_cast<size_t>(D) * H * W * sizeof(float));
  const float s2 = 2.f * loss_weight * loss_grad;
  for (int v = 0; v < views; ++v) {
    const Camera cam = make_camera(m16, v16 + 16 * v, p16 + 16 * v);
#pragma omp parallel
    {
      std::vector<float> sig, dpre;
#pragma omp for collapse(2) schedule(dynamic, 4)
      for (int py = 0; py < OH; ++py)
        for (int px = 0; px < OW; ++px) {
          const Ray r = make_ray(cam, d, px, py, OW, OH, OD);
          if (r.steps <= 0) continue;
          const int64_t pix = (static_cast<int64_t>(v) * OH + py) * OW + px;
          float gL[3];
          for (int c = 0; c < 3; ++c) gL[c] = s2 * (images[pix * 4 + c] - target[pix * 3 + c]);
          sig.resize(r.steps);
          dpre.resize(r.steps + 1);
          dpre[0] = 0.f;
          for (int k = 0; k < r.steps; ++k) {
            const float kf = static_cast<float>(k);
            sig[k] = depth_step * sample_linear(rho, d, std::fma(kf, r.dx, r.sx) - 0.5f,
                                                std::fma(kf, r.dy, r.sy) - 0.5f, std::fma(kf, r.dz, r.sz) - 0.5f);
            dpre[k + 1] = std::max(dpre[k] + sig[k], 0.f);
          }
          const float lg = light_rgb[0] * gL[0] + light_rgb[1] * gL[1] + light_rgb[2] * gL[2];
          float ga = 0.f;  // gradient of the optical depth accumulator
          for (int k = r.steps - 1; k >= 0; --k) {
            const float e = std::exp(-dpre[k + 1]);
            const float gd = ga - e * sig[k] * lg;
            // cell (rgb * s, s) * depth_step -> density
            const float grho = depth_step * (e * lg + gd);
            const float kf = static_cast<float>(k);
            scatter_linear(g_rho, d, std::fma(kf, r.dx, r.sx) - 0.5f, std::fma(kf, r.dy, r.sy) - 0.5f,
                           std::fma(kf, r.dz, r.sz) - 0.5f, grho);
            ga = gd;
          }
        }
    }
  }
}

}  // extern "C"
