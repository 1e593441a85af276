// Physics of one env, written for one 64-lane wavefront per env.
//
// Every stage is a lane-strided loop over independent work items (bodies of
// one tree level, dofs, pairs, rows, matrix entries) followed by KSYNC().
// Values that live across a KSYNC() are either in the LDS workspace or
// wave-uniform (results of wave reductions), so the same source also runs
// with WAVE == 1 as a serial interpreter (MJ_EMU builds; test-only).
//
// Stage order per substep follows the reference pipeline
// (reference src/batched.cpp:980-1228: kinematics, forward, collision, solver,
// Euler) with MuJoCo semantics where the reference departs from them
// (multi-contact narrow phase, warm start, MuJoCo termination test).
#pragma once
#include <math.h>
#include <stdint.h>

#include "dmodel.h"

// MJK_BIG 0 compiles the step for models of at most SPARSE_ROWS_NV (32) dofs
// only: the paths of larger models (compressed constraint rows, and beyond 64
// dofs tree-sparse products, streamed J^T f and memory-resident triangular
// solves) drop out of the kernel, and with them their registers. The default
// builds every path.
#ifndef MJK_BIG
#define MJK_BIG 1
#endif

// MJK_HALF 1 compiles the step with 32 lanes per env: each 64-lane wave
// steps two envs, one per half (lanes 0-31 and 32-63), each with its own
// workspace; only for models of at most 32 dofs (MJK_BIG 0).
#if defined(MJ_EMU)
#define KDEV static inline
#define WAVE 1
#define LANE 0
#define KSYNC() ((void)0)
#elif defined(MJK_HALF) && MJK_HALF
#if MJK_BIG
#error "the 32-lane step is built for models of at most 32 dofs only (MJK_BIG 0)"
#endif
#define KDEV __device__ __forceinline__
#define WAVE 32
#define LANE ((int)(threadIdx.x & 31))
#define KSYNC() __syncthreads()
#else
#define KDEV __device__ __forceinline__
#define WAVE 64
#define LANE ((int)threadIdx.x)
#define KSYNC() __syncthreads()
#endif

namespace mjk {

// Per-stage cycle accounting, compiled in only for profiling builds
// (MJK_PROFILE); the product build has empty marks.
enum { KP_KIN = 0, KP_COMPOS, KP_CRB, KP_FACTOR, KP_COMVEL, KP_SMOOTH, KP_COLL, KP_CONSTR, KP_SOLVE_INIT, KP_LS,
       KP_COSTGRAD, KP_DIR, KP_SOLVE_REST, KP_INTEGRATE, KP_IO, KP_SPARE, KP_NSOLVE, KP_N };
struct KProfAcc {
  unsigned long long t[KP_N];
  unsigned long long last;
};
#if defined(MJK_PROFILE) && !defined(MJ_EMU)
#define KPROF_START(pa) ((pa)->last = clock64())
#define KPROF_MARK(pa, i)                          \
  do {                                             \
    unsigned long long t_ = clock64();             \
    (pa)->t[i] += t_ - (pa)->last;                 \
    (pa)->last = t_;                               \
  } while (0)
#else
#define KPROF_START(pa) ((void)0)
#define KPROF_MARK(pa, i) ((void)(pa))
#endif

#define KMINVAL 1e-15f
#define KMINIMP 0.0001f
#define KMAXIMP 0.9999f

// The model and the workspace layout of one sim, as one device-memory copy
// the step kernels read through a pointer.
struct KParams {
  DModel m;
  Layout L;
};

// kparams(): the kernels' view of that copy as constant memory (address
// space 4: loads through it are scalar loads the compiler may repeat).
// kfresh(x): the same object, for x inside such a copy, through a pointer the
// compiler cannot prove equal to the one before: the fields a stage reads are
// loaded again at the stage instead of staying live in scalar registers
// across the whole step, where they spill into vector-register lanes and
// every use reloads one with a v_readlane. x must live in device memory (a
// kernel's by-value argument does not: its address is a private copy).
#if defined(MJ_EMU)
static inline const KParams& kparams(const KParams* p) { return *p; }
template <class T>
static inline const T& kfresh(const T& x) { return x; }
#else
__device__ __forceinline__ const KParams& kparams(const KParams* p) {
  typedef const __attribute__((address_space(4))) KParams* cptr;
  return *(const KParams*)(cptr)p;
}
template <class T>
__device__ __forceinline__ const T& kfresh(const T& x) {
  typedef const __attribute__((address_space(4))) T* cptr;
  cptr p = (cptr)&x;
  asm volatile("" : "+s"(p));
  return *(const T*)p;
}
#endif

// bit casts between int and float (row type bits kept in float slots)
KDEV float ki2f(int v) { return __builtin_bit_cast(float, v); }
KDEV int kf2i(float v) { return __builtin_bit_cast(int, v); }

// ------------------------------------------------------------------ wave helpers
// Wave-wide reductions: DPP exchanges inside each 16-lane row (quad
// permutations, half-row and row mirrors) leave every lane of a row holding
// the row's total r_k; row broadcasts then add row 0's total into row 1 and
// row 2's into row 3 (row_bcast:15), and row 1's into row 3 (row_bcast:31),
// and lane 63's (r2 + r3) + (r0 + r1) is read into a scalar register: every
// lane gets the same bit pattern and the result can drive uniform control
// flow. With 32 lanes per env each half of the wave stops after the first
// broadcast and a lane takes its half's total (lane 31 or lane 63).
#if WAVE >= 32
KDEV int dpp_i(int v, int ctrl_sel) {
  switch (ctrl_sel) {
    case 0: return __builtin_amdgcn_update_dpp(0, v, 0xB1, 0xF, 0xF, false);   // quad_perm [1,0,3,2]
    case 1: return __builtin_amdgcn_update_dpp(0, v, 0x4E, 0xF, 0xF, false);   // quad_perm [2,3,0,1]
    case 2: return __builtin_amdgcn_update_dpp(0, v, 0x141, 0xF, 0xF, false);  // row_half_mirror
    default: return __builtin_amdgcn_update_dpp(0, v, 0x140, 0xF, 0xF, false); // row_mirror
  }
}
// row broadcasts into the rows of row_mask; the other rows return `old`
KDEV int dpp_bc15(int old, int v) { return __builtin_amdgcn_update_dpp(old, v, 0x142, 0xA, 0xF, false); }
KDEV int dpp_bc31(int old, int v) { return __builtin_amdgcn_update_dpp(old, v, 0x143, 0xC, 0xF, false); }
KDEV float dpp_f(float v, int c) { return __int_as_float(dpp_i(__float_as_int(v), c)); }
KDEV int rl_abs(int v, int lane) { return __builtin_amdgcn_readlane(v, lane); }
KDEV float rl_absf(float v, int lane) { return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), lane)); }
KDEV double rl_absd(double v, int lane) {
  long long b = __double_as_longlong(v);
  int lo = __builtin_amdgcn_readlane((int)(b & 0xffffffffll), lane);
  int hi = __builtin_amdgcn_readlane((int)(b >> 32), lane);
  return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}
KDEV double dpp_bcd(double v, bool b31) {
  long long b = __double_as_longlong(v);
  int lo = (int)(b & 0xffffffffll), hi = (int)(b >> 32);
  lo = b31 ? dpp_bc31(0, lo) : dpp_bc15(0, lo);
  hi = b31 ? dpp_bc31(0, hi) : dpp_bc15(0, hi);
  return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}
#if WAVE == 64
// value of lane `lane` (uniform) of this env's lanes
KDEV float rl_f(float v, int lane) { return rl_absf(v, lane); }
KDEV double rl_d(double v, int lane) { return rl_absd(v, lane); }
#define WHALF_SEL(a, b) (a)
#else
KDEV bool whi() { return (threadIdx.x & 32) != 0; }
KDEV float rl_f(float v, int lane) { float a = rl_absf(v, lane), b = rl_absf(v, lane + 32); return whi() ? b : a; }
KDEV double rl_d(double v, int lane) { double a = rl_absd(v, lane), b = rl_absd(v, lane + 32); return whi() ? b : a; }
#define WHALF_SEL(a, b) (whi() ? (b) : (a))
#endif
#endif
KDEV float wsum(float v) {
#if WAVE >= 32
  v += dpp_f(v, 0);
  v += dpp_f(v, 1);
  v += dpp_f(v, 2);
  v += dpp_f(v, 3);
  v += __int_as_float(dpp_bc15(0, __float_as_int(v)));
#if WAVE == 64
  v += __int_as_float(dpp_bc31(0, __float_as_int(v)));
  return rl_absf(v, 63);
#else
  return WHALF_SEL(rl_absf(v, 31), rl_absf(v, 63));
#endif
#else
  return v;
#endif
}
KDEV double wsumd(double v) {
#if WAVE >= 32
  for (int c = 0; c < 4; c++) {
    long long b = __double_as_longlong(v);
    int lo = dpp_i((int)(b & 0xffffffffll), c), hi = dpp_i((int)(b >> 32), c);
    v += __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
  }
  v += dpp_bcd(v, false);
#if WAVE == 64
  v += dpp_bcd(v, true);
  return rl_absd(v, 63);
#else
  return WHALF_SEL(rl_absd(v, 31), rl_absd(v, 63));
#endif
#else
  return v;
#endif
}
KDEV int wsumi(int v) {
#if WAVE >= 32
  for (int c = 0; c < 4; c++) v += dpp_i(v, c);
  v += dpp_bc15(0, v);
#if WAVE == 64
  v += dpp_bc31(0, v);
  return rl_abs(v, 63);
#else
  return WHALF_SEL(rl_abs(v, 31), rl_abs(v, 63));
#endif
#else
  return v;
#endif
}
KDEV int wori(int v) {
#if WAVE >= 32
  for (int c = 0; c < 4; c++) v |= dpp_i(v, c);
  v |= dpp_bc15(0, v);
#if WAVE == 64
  v |= dpp_bc31(0, v);
  return rl_abs(v, 63);
#else
  return WHALF_SEL(rl_abs(v, 31), rl_abs(v, 63));
#endif
#else
  return v;
#endif
}
KDEV int wmini(int v) {
#if WAVE >= 32
  for (int c = 0; c < 4; c++) { int t = dpp_i(v, c); v = t < v ? t : v; }
  { int t = dpp_bc15(v, v); v = t < v ? t : v; }  // rows outside the mask compare with themselves
#if WAVE == 64
  { int t = dpp_bc31(v, v); v = t < v ? t : v; }
  return rl_abs(v, 63);
#else
  return WHALF_SEL(rl_abs(v, 31), rl_abs(v, 63));
#endif
#else
  return v;
#endif
}
// exclusive prefix sum over this env's lanes
KDEV int wscan_excl(int v) {
#if WAVE >= 32
  int x = v;
  for (int o = 1; o < WAVE; o <<= 1) {
    int t = __shfl_up(x, o, WAVE);
    if (LANE >= o) x += t;
  }
  return x - v;
#else
  (void)v;
  return 0;
#endif
}

KDEV int* iptr(float* w, int off) { return reinterpret_cast<int*>(w + off); }
KDEV const int* iptr(const float* w, int off) { return reinterpret_cast<const int*>(w + off); }
KDEV double* dptr(float* w, int off) { return reinterpret_cast<double*>(w + off); }

// (row, col) with col <= row of lower-triangular entry t (row-major packing)
// 1 / x in FP64 for x > 0 (not denormal): the hardware reciprocal refined by
// two Newton-Raphson steps (within an ulp), without the scaling and fix-up
// steps of an IEEE division
KDEV double krcp(double x) {
#if defined(MJ_EMU)
  return 1.0 / x;
#else
  double r = __builtin_amdgcn_rcp(x);
  r = fma(fma(-x, r, 1.0), r, r);
  return fma(fma(-x, r, 1.0), r, r);
#endif
}

// entry (i, j), j <= i, of a lower triangle packed row by row (the Newton
// Hessian and its factor)
KDEV int hx(int i, int j) { return ((i * (i + 1)) >> 1) + j; }

KDEV void tri_rc(int t, int* r, int* c) {
  int row = (int)((sqrtf(8.0f * (float)t + 1.0f) - 1.0f) * 0.5f);
  while ((row + 1) * (row + 2) / 2 <= t) row++;
  while (row * (row + 1) / 2 > t) row--;
  *r = row;
  *c = t - row * (row + 1) / 2;
}

// ------------------------------------------------------------------ small math
KDEV void quat2mat(const float* q, float* R) {
  float w = q[0], x = q[1], y = q[2], z = q[3];
  R[0] = w * w + x * x - y * y - z * z; R[1] = 2.f * (x * y - w * z); R[2] = 2.f * (x * z + w * y);
  R[3] = 2.f * (x * y + w * z); R[4] = w * w - x * x + y * y - z * z; R[5] = 2.f * (y * z - w * x);
  R[6] = 2.f * (x * z - w * y); R[7] = 2.f * (y * z + w * x); R[8] = w * w - x * x - y * y + z * z;
}
KDEV void mulquat(const float* a, const float* b, float* o) {
  float r0 = a[0] * b[0] - a[1] * b[1] - a[2] * b[2] - a[3] * b[3];
  float r1 = a[0] * b[1] + a[1] * b[0] + a[2] * b[3] - a[3] * b[2];
  float r2 = a[0] * b[2] - a[1] * b[3] + a[2] * b[0] + a[3] * b[1];
  float r3 = a[0] * b[3] + a[1] * b[2] - a[2] * b[1] + a[3] * b[0];
  o[0] = r0; o[1] = r1; o[2] = r2; o[3] = r3;
}
KDEV void normq(float* q) {
  float n = sqrtf(q[0] * q[0] + q[1] * q[1] + q[2] * q[2] + q[3] * q[3]);
  if (n < KMINVAL) { q[0] = 1.f; q[1] = q[2] = q[3] = 0.f; return; }
  float in = 1.0f / n;
  q[0] *= in; q[1] *= in; q[2] *= in; q[3] *= in;
}
KDEV void mv3(const float* R, const float* v, float* o) {
  float a = R[0] * v[0] + R[1] * v[1] + R[2] * v[2];
  float b = R[3] * v[0] + R[4] * v[1] + R[5] * v[2];
  float c = R[6] * v[0] + R[7] * v[1] + R[8] * v[2];
  o[0] = a; o[1] = b; o[2] = c;
}
KDEV void mtv3(const float* R, const float* v, float* o) {
  float a = R[0] * v[0] + R[3] * v[1] + R[6] * v[2];
  float b = R[1] * v[0] + R[4] * v[1] + R[7] * v[2];
  float c = R[2] * v[0] + R[5] * v[1] + R[8] * v[2];
  o[0] = a; o[1] = b; o[2] = c;
}
KDEV void cross3(const float* a, const float* b, float* o) {
  float x = a[1] * b[2] - a[2] * b[1], y = a[2] * b[0] - a[0] * b[2], z = a[0] * b[1] - a[1] * b[0];
  o[0] = x; o[1] = y; o[2] = z;
}
KDEV float dot3(const float* a, const float* b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }
KDEV float clampf(float x, float lo, float hi) { return x < lo ? lo : (x > hi ? hi : x); }

// sum_k a[k*sa] * b[k*sb]: four independent partial sums so that the LDS
// loads of consecutive terms overlap instead of serialising on one chain
KDEV float kdot(const float* a, int sa, const float* b, int sb, int n) {
  float s0 = 0.f, s1 = 0.f, s2 = 0.f, s3 = 0.f;
  int k = 0;
  for (; k + 3 < n; k += 4) {
    s0 += a[k * sa] * b[k * sb];
    s1 += a[(k + 1) * sa] * b[(k + 1) * sb];
    s2 += a[(k + 2) * sa] * b[(k + 2) * sb];
    s3 += a[(k + 3) * sa] * b[(k + 3) * sb];
  }
  for (; k < n; k++) s0 += a[k * sa] * b[k * sb];
  return (s0 + s1) + (s2 + s3);
}
// sum_k c[k] * a[k*sa] * b[k*sb]
KDEV float kdot3(const float* c, const float* a, int sa, const float* b, int sb, int n) {
  float s0 = 0.f, s1 = 0.f, s2 = 0.f, s3 = 0.f;
  int k = 0;
  for (; k + 3 < n; k += 4) {
    s0 += c[k] * a[k * sa] * b[k * sb];
    s1 += c[k + 1] * a[(k + 1) * sa] * b[(k + 1) * sb];
    s2 += c[k + 2] * a[(k + 2) * sa] * b[(k + 2) * sb];
    s3 += c[k + 3] * a[(k + 3) * sa] * b[(k + 3) * sb];
  }
  for (; k < n; k++) s0 += c[k] * a[k * sa] * b[k * sb];
  return (s0 + s1) + (s2 + s3);
}

KDEV void mul_inert(const float* i, const float* v, float* r) {
  r[0] = i[0] * v[0] + i[3] * v[1] + i[4] * v[2] - i[8] * v[4] + i[7] * v[5];
  r[1] = i[3] * v[0] + i[1] * v[1] + i[5] * v[2] + i[8] * v[3] - i[6] * v[5];
  r[2] = i[4] * v[0] + i[5] * v[1] + i[2] * v[2] - i[7] * v[3] + i[6] * v[4];
  r[3] = i[8] * v[1] - i[7] * v[2] + i[9] * v[3];
  r[4] = i[6] * v[2] - i[8] * v[0] + i[9] * v[4];
  r[5] = i[7] * v[0] - i[6] * v[1] + i[9] * v[5];
}
KDEV void cross_motion(const float* v, const float* u, float* r) {
  float a[3], b[3], c[3];
  cross3(v, u, a); cross3(v, u + 3, b); cross3(v + 3, u, c);
  r[0] = a[0]; r[1] = a[1]; r[2] = a[2];
  r[3] = b[0] + c[0]; r[4] = b[1] + c[1]; r[5] = b[2] + c[2];
}
KDEV void cross_force(const float* v, const float* f, float* r) {
  float a[3], b[3], c[3];
  cross3(v, f, a); cross3(v + 3, f + 3, b); cross3(v, f + 3, c);
  r[0] = a[0] + b[0]; r[1] = a[1] + b[1]; r[2] = a[2] + b[2];
  r[3] = c[0]; r[4] = c[1]; r[5] = c[2];
}

// ------------------------------------------------------------------ kinematics
// mj_kinematics, level by level (bodies of one tree depth in parallel).
// Reference src/compat/mkx_kernels/kinematics.hpp:79-144.
KDEV void st_kinematics(const DModel& m, const Layout& L, float* w) {
  float *xpos = w + L.xpos, *xquat = w + L.xquat, *xmat = w + L.xmat, *xipos = w + L.xipos;
  float *xanc = w + L.xanchor, *xax = w + L.xaxis;
  const float* qpos = w + L.qpos;
  if (LANE == 0) {
    xpos[0] = xpos[1] = xpos[2] = 0.f;
    xquat[0] = 1.f; xquat[1] = xquat[2] = xquat[3] = 0.f;
    for (int i = 0; i < 9; i++) xmat[i] = (i % 4 == 0) ? 1.f : 0.f;
    xipos[0] = xipos[1] = xipos[2] = 0.f;
  }
  KSYNC();
  // pass 1, all bodies at once: each body's frame relative to its parent's
  // (free joints: in the world), kept in its own xpos / xquat, and its
  // joints' anchors and axes in the parent's frame (world for free joints),
  // kept in xanchor / xaxis; the model reads and the joints' trigonometry
  // leave the level-by-level chain
  for (int b = 1 + LANE; b < m.nbody; b += WAVE) {
    int ja = m.body_jntadr[b], jn = m.body_jntnum[b];
    float pos[3], q[4];
    if (jn > 0 && m.jnt_type[ja] == JFREE) {
      int qa = m.jnt_qposadr[ja];
      pos[0] = qpos[qa]; pos[1] = qpos[qa + 1]; pos[2] = qpos[qa + 2];
      q[0] = qpos[qa + 3]; q[1] = qpos[qa + 4]; q[2] = qpos[qa + 5]; q[3] = qpos[qa + 6];
      normq(q);
      for (int k = 0; k < 3; k++) { xanc[3 * ja + k] = pos[k]; xax[3 * ja + k] = m.jnt_axis[3 * ja + k]; }
    } else {
      for (int k = 0; k < 3; k++) pos[k] = m.body_pos[3 * b + k];
      for (int k = 0; k < 4; k++) q[k] = m.body_quat[4 * b + k];
      for (int j = ja; j < ja + jn; j++) {
        int qa = m.jnt_qposadr[j];
        float R[9], anc[3], ax[3];
        quat2mat(q, R);
        mv3(R, m.jnt_pos + 3 * j, anc);
        anc[0] += pos[0]; anc[1] += pos[1]; anc[2] += pos[2];
        mv3(R, m.jnt_axis + 3 * j, ax);
        for (int k = 0; k < 3; k++) { xanc[3 * j + k] = anc[k]; xax[3 * j + k] = ax[k]; }
        int jt = m.jnt_type[j];
        if (jt == JSLIDE) {
          float sl = qpos[qa] - m.qpos0[qa];
          for (int k = 0; k < 3; k++) pos[k] += ax[k] * sl;
        } else {
          float ql[4];
          if (jt == JHINGE) {
            float ang = 0.5f * (qpos[qa] - m.qpos0[qa]);
            float sn = sinf(ang), cs = cosf(ang);
            ql[0] = cs; ql[1] = m.jnt_axis[3 * j] * sn; ql[2] = m.jnt_axis[3 * j + 1] * sn; ql[3] = m.jnt_axis[3 * j + 2] * sn;
          } else {
            ql[0] = qpos[qa]; ql[1] = qpos[qa + 1]; ql[2] = qpos[qa + 2]; ql[3] = qpos[qa + 3];
            normq(ql);
          }
          mulquat(q, ql, q);
          float R2[9], r[3];
          quat2mat(q, R2);
          mv3(R2, m.jnt_pos + 3 * j, r);
          for (int k = 0; k < 3; k++) pos[k] = anc[k] - r[k];
        }
      }
    }
    for (int k = 0; k < 3; k++) xpos[3 * b + k] = pos[k];
    for (int k = 0; k < 4; k++) xquat[4 * b + k] = q[k];
  }
  KSYNC();
  // pass 2, level by level: compose with the parent's world frame
  for (int lv = 0; lv < m.nlevel; lv++) {
    int a = m.level_adr[lv], n = m.level_adr[lv + 1] - a;
    for (int i = LANE; i < n; i += WAVE) {
      int b = m.level_body[a + i];
      int par = m.body_parentid[b];
      int ja = m.body_jntadr[b], jn = m.body_jntnum[b];
      float pos[3], q[4];
      for (int k = 0; k < 3; k++) pos[k] = xpos[3 * b + k];
      for (int k = 0; k < 4; k++) q[k] = xquat[4 * b + k];
      if (!(jn > 0 && m.jnt_type[ja] == JFREE)) {
        const float *Rp = xmat + 9 * par, *pp = xpos + 3 * par;
        float t[3];
        mv3(Rp, pos, t);
        pos[0] = pp[0] + t[0]; pos[1] = pp[1] + t[1]; pos[2] = pp[2] + t[2];
        mulquat(xquat + 4 * par, q, q);
        for (int j = ja; j < ja + jn; j++) {
          float an[3], ax[3];
          mv3(Rp, xanc + 3 * j, an);
          mv3(Rp, xax + 3 * j, ax);
          for (int k = 0; k < 3; k++) { xanc[3 * j + k] = pp[k] + an[k]; xax[3 * j + k] = ax[k]; }
        }
        normq(q);
      }
      for (int k = 0; k < 3; k++) xpos[3 * b + k] = pos[k];
      for (int k = 0; k < 4; k++) xquat[4 * b + k] = q[k];
      float R[9];
      quat2mat(q, R);
      for (int k = 0; k < 9; k++) xmat[9 * b + k] = R[k];
      float t[3];
      if (L.rand >= 0) {  // randomised COM: the body-frame shift moves the inertial frame
        const float* sh = w + L.rand + rand_com(m.nbody) + 3 * b;
        float ip[3] = {m.body_ipos[3 * b] + sh[0], m.body_ipos[3 * b + 1] + sh[1], m.body_ipos[3 * b + 2] + sh[2]};
        mv3(R, ip, t);
      } else {
        mv3(R, m.body_ipos + 3 * b, t);
      }
      for (int k = 0; k < 3; k++) xipos[3 * b + k] = pos[k] + t[k];
    }
    KSYNC();
  }
  float *gx = w + L.gxpos, *gm = w + L.gxmat;
  for (int g = LANE; g < m.ngeom; g += WAVE) {
    int b = m.geom_bodyid[g];
    float t[3], qg[4];
    mv3(xmat + 9 * b, m.geom_pos + 3 * g, t);
    for (int k = 0; k < 3; k++) gx[3 * g + k] = xpos[3 * b + k] + t[k];
    mulquat(xquat + 4 * b, m.geom_quat + 4 * g, qg);
    quat2mat(qg, gm + 9 * g);
  }
  KSYNC();
}

// mj_comPos: subtree COM (bodies are numbered in depth-first order, so a
// subtree is the id range [b, body_subtree_end[b])), cinert and cdof.
// Reference forward.hpp:146-233.
KDEV void st_com_pos(const DModel& m, const Layout& L, float* w) {
  const float *xipos = w + L.xipos, *xquat = w + L.xquat, *xmat = w + L.xmat;
  float *scom = w + L.scom, *cinert = w + L.cinert, *cdof = w + L.cdof;
  const float* msc = L.rand >= 0 ? w + L.rand : nullptr;  // randomised mass scales (mass and inertia)
  for (int b = LANE; b < m.nbody; b += WAVE) {
    float s[3] = {0.f, 0.f, 0.f}, sm = 0.f;
    for (int k = b; k < m.body_subtree_end[b]; k++) {
      float mk = msc ? m.body_mass[k] * msc[k] : m.body_mass[k];
      s[0] += mk * xipos[3 * k]; s[1] += mk * xipos[3 * k + 1]; s[2] += mk * xipos[3 * k + 2];
      sm += mk;
    }
    if (!msc) sm = m.body_subtreemass[b];
    for (int k = 0; k < 3; k++) scom[3 * b + k] = sm < KMINVAL ? xipos[3 * b + k] : s[k] / sm;
  }
  KSYNC();
  for (int b = LANE; b < m.nbody; b += WAVE) {
    float* o = cinert + 10 * b;
    if (b == 0) { for (int k = 0; k < 10; k++) o[k] = 0.f; continue; }
    float qi[4], R[9];
    mulquat(xquat + 4 * b, m.body_iquat + 4 * b, qi);
    quat2mat(qi, R);
    float sc = msc ? msc[b] : 1.f;
    const float I[3] = {m.body_inertia[3 * b] * sc, m.body_inertia[3 * b + 1] * sc, m.body_inertia[3 * b + 2] * sc};
    float mass = m.body_mass[b] * sc;
    const float* c = scom + 3 * m.body_rootid[b];
    float d[3] = {xipos[3 * b] - c[0], xipos[3 * b + 1] - c[1], xipos[3 * b + 2] - c[2]};
    float A[9];
    for (int r = 0; r < 3; r++)
      for (int k = 0; k < 3; k++) A[3 * r + k] = R[3 * r] * I[0] * R[3 * k] + R[3 * r + 1] * I[1] * R[3 * k + 1] + R[3 * r + 2] * I[2] * R[3 * k + 2];
    float sq = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
    o[0] = A[0] + mass * (sq - d[0] * d[0]);
    o[1] = A[4] + mass * (sq - d[1] * d[1]);
    o[2] = A[8] + mass * (sq - d[2] * d[2]);
    o[3] = A[1] - mass * d[0] * d[1];
    o[4] = A[2] - mass * d[0] * d[2];
    o[5] = A[5] - mass * d[1] * d[2];
    o[6] = mass * d[0]; o[7] = mass * d[1]; o[8] = mass * d[2]; o[9] = mass;
  }
  const float *xanc = w + L.xanchor, *xax = w + L.xaxis;
  for (int k = LANE; k < m.nv; k += WAVE) {
    int j = m.dof_jntid[k], b = m.dof_bodyid[k];
    int jt = m.jnt_type[j], i = k - m.jnt_dofadr[j];
    float* cd = cdof + 6 * k;
    const float* c = scom + 3 * m.body_rootid[b];
    const float* anc = xanc + 3 * j;
    float off[3] = {c[0] - anc[0], c[1] - anc[1], c[2] - anc[2]};
    float ax[3];
    bool rot = true;
    if (jt == JFREE && i < 3) {
      for (int q = 0; q < 6; q++) cd[q] = 0.f;
      cd[3 + i] = 1.f;
      rot = false;
    } else if (jt == JFREE || jt == JBALL) {
      int col = jt == JFREE ? i - 3 : i;
      ax[0] = xmat[9 * b + col]; ax[1] = xmat[9 * b + 3 + col]; ax[2] = xmat[9 * b + 6 + col];
    } else if (jt == JSLIDE) {
      cd[0] = cd[1] = cd[2] = 0.f;
      cd[3] = xax[3 * j]; cd[4] = xax[3 * j + 1]; cd[5] = xax[3 * j + 2];
      rot = false;
    } else {
      ax[0] = xax[3 * j]; ax[1] = xax[3 * j + 1]; ax[2] = xax[3 * j + 2];
    }
    if (rot) {
      cd[0] = ax[0]; cd[1] = ax[1]; cd[2] = ax[2];
      cross3(ax, off, cd + 3);
    }
  }
  KSYNC();
}

// qM / qLD storage: compact rows (DModel::nM entries) beyond COMPACT_M_NV
// dofs, else dense nv x nv
KDEV bool m_compact(const DModel& m) { return MJK_BIG && m.nv > COMPACT_M_NV; }
// index of entry (i, j), j == i or a dof ancestor of i (any j when dense)
KDEV int mxi(const DModel& m, int i, int j) {
  return m_compact(m) ? m.dof_madr[i] + m.dof_depth[i] - m.dof_depth[j] : i * m.nv + j;
}

// entry (i, j), j <= i, of qM in either storage (zero off the dof tree)
KDEV float m_entry(const DModel& m, const float* M, int i, int j) {
  if (!m_compact(m)) return M[i * m.nv + j];
  return i < m.dof_subtree_end[j] ? M[mxi(m, i, j)] : 0.f;  // j is i or an ancestor
}

// mj_crb: composite inertia by subtree ranges, then qM entries along each dof's
// ancestor chain (reference forward.hpp:240-279).
KDEV void st_crb(const DModel& m, const Layout& L, float* w) {
  const float *cinert = w + L.cinert, *cdof = w + L.cdof;
  float *crb = w + L.crb, *buf = w + L.buf6, *M = w + L.qM;
  int nv = m.nv;
  // one lane per (body, component): the subtree sums run side by side
  for (int t = LANE; t < 10 * m.nbody; t += WAVE) {
    int b = t / 10, q = t - 10 * b;
    float s = 0.f;
    if (b > 0)
      for (int k = b; k < m.body_subtree_end[b]; k++) s += cinert[10 * k + q];
    crb[t] = s;
  }
  bool compact = m_compact(m);
  if (!compact)  // compact rows hold no zeros: every entry is written below
    for (int t = LANE; t < nv * nv; t += WAVE) M[t] = 0.f;
  KSYNC();
  for (int i = LANE; i < nv; i += WAVE) mul_inert(crb + 10 * m.dof_bodyid[i], cdof + 6 * i, buf + 6 * i);
  KSYNC();
  // each dof's ancestors from its chain table (no dependent parent loads)
  for (int i = LANE; i < nv; i += WAVE) {
    const float* bi = buf + 6 * i;
    const MJK_G int* ch = m.dof_chain + i * MAX_DOF_DEPTH;
    for (int a = 0, n = m.dof_nchain[i]; a < n; a++) {
      int j = ch[a];
      const float* cj = cdof + 6 * j;
      float s = cj[0] * bi[0] + cj[1] * bi[1] + cj[2] * bi[2] + cj[3] * bi[3] + cj[4] * bi[4] + cj[5] * bi[5];
      if (compact) {
        M[m.dof_madr[i] + a] = s;  // the chain's a-th entry is the row's a-th
      } else {
        M[i * nv + j] = s;
        M[j * nv + i] = s;
      }
    }
    M[mxi(m, i, i)] += m.dof_armature[i];
  }
  KSYNC();
}

// Tree-sparse L^T D L factorisation (MuJoCo mj_factorM) of A (dense storage)
// into LD. Step k updates the entries (i, j), i and j strict ancestors of k,
// j an ancestor of i, in parallel, from the still unscaled row k (row k is
// final once step k starts: its entries are only touched by steps of its
// descendants, which have larger indices). Rows are scaled by 1/D in one
// final pass. Returns ST_M_NOT_PD on a non-positive pivot.
// DENSE: the dense elimination below is compiled in (not in the 168-VGPR
// kernels, whose register allocation it upsets: quadruped +0.7 %)
template <bool DENSE = true>
KDEV int factor_ld(const DModel& m, const float* A, float* LD) {
  int nv = m.nv;
  // beyond 64 dofs only the tree's entries are copied and scaled (row k:
  // its diagonal and its ancestors), all its readers touch; up to 64 the
  // dense matrix, whose zeros the register solve reads
  bool sparse = m_compact(m);  // compact rows: every stored entry is a tree entry
  if (sparse) {
    for (int t = LANE; t < m.nM; t += WAVE) LD[t] = A[t];
  } else {
    for (int t = LANE; t < nv * nv; t += WAVE) LD[t] = A[t];
  }
  KSYNC();
  int bad = 0;
  // trees of at most 32 dofs with more levels than half their dofs (long
  // chains: the humanoid's 14 levels for 27 dofs): dense elimination, whose
  // nv / 2 phases need no table loads (humanoid -2.1 % per control step);
  // the others (the quadruped's 8 levels for 18 dofs, free bodies' 6-dof
  // chains): the level tables
  bool dense = DENSE && !sparse && nv <= SPARSE_ROWS_NV && 2 * m.nflev > nv;
  if (dense) {
    // Dense elimination from the last dof to the first, two dofs (k, k - 1)
    // per LDS phase: with c = A[k][k-1] / A[k][k] and the pair's second pivot
    // d1 = A[k-1][k-1] - c A[k][k-1], the leading entries (i, j),
    // j <= i < k - 1, take both rank-1 updates at once,
    //   A[i][j] -= A[k][i] A[k][j] / A[k][k] + b_i b_j / d1,  b_x = A[k-1][x] - c A[k][x];
    // row k - 1 takes its own update in one pass at the end. Entries outside
    // the dof tree stay zero (A[k][i] = 0 unless i is an ancestor of k), so
    // this is the tree-sparse factor, without the level tables' loads. A
    // lane's entries of a leading triangle, enumerated row-major, are the
    // same at every step: for up to TR rounds of lanes they are computed once.
    constexpr int TR = WAVE >= 32 ? 3 * (64 / WAVE) : 1;
    int tri_ij[TR];
    int ntri0 = nv >= 2 ? (nv - 2) * (nv - 1) / 2 : 0;
    bool cached = ntri0 <= TR * WAVE;
#pragma unroll
    for (int r = 0; r < TR; r++) {
      int t = LANE + r * WAVE, a = 0, b = 0;
      if (cached && t < ntri0) tri_rc(t, &a, &b);
      tri_ij[r] = a | (b << 16);
    }
    for (int k = nv - 1; k >= 1; k -= 2) {
      float d0 = LD[k * nv + k], h = LD[k * nv + k - 1], a11 = LD[(k - 1) * nv + k - 1];
      if (!(d0 > KMINVAL)) { bad = 1; d0 = KMINVAL; }
      float i0 = 1.f / d0, c = h * i0, d1 = a11 - c * h;
      if (!(d1 > KMINVAL)) { bad = 1; d1 = KMINVAL; }
      float i1 = 1.f / d1;
      int nt = (k - 1) * k / 2;  // entries (i, j), j <= i < k - 1
      const float *rk = LD + k * nv, *rk1 = LD + (k - 1) * nv;
      if (cached) {
#pragma unroll
        for (int r = 0; r < TR; r++) {
          if (LANE + r * WAVE >= nt) continue;
          int i = tri_ij[r] & 0xffff, j = tri_ij[r] >> 16;
          float ai = rk[i], aj = rk[j];
          float bi = rk1[i] - c * ai, bj = rk1[j] - c * aj;
          LD[i * nv + j] -= ai * aj * i0 + bi * bj * i1;
        }
      } else {
        for (int t = LANE; t < nt; t += WAVE) {
          int i, j;
          tri_rc(t, &i, &j);
          float ai = rk[i], aj = rk[j];
          float bi = rk1[i] - c * ai, bj = rk1[j] - c * aj;
          LD[i * nv + j] -= ai * aj * i0 + bi * bj * i1;
        }
      }
      KSYNC();
    }
    // the pairs' second rows (their diagonal included), each from its first
    // row, which is final: rows of different pairs do not read each other
    for (int k = nv - 1; k >= 1; k -= 2) {
      float d0 = LD[k * nv + k];
      if (!(d0 > KMINVAL)) d0 = KMINVAL;
      float c = LD[k * nv + k - 1] * (1.f / d0);
      for (int x = LANE; x <= k - 1; x += WAVE) LD[(k - 1) * nv + x] -= c * LD[k * nv + x];
    }
    KSYNC();
  }
  // dofs of one tree level are neither ancestors nor descendants of each
  // other, so their steps run together: each entry (i, j) of the level's list
  // sums the contributions of the level's dofs that have both i and j as
  // strict ancestors. Levels go deepest first, so a dof's row is final when read.
  for (int l = 0; !dense && l < m.nflev; l++) {
    for (int e = m.flev_adr[l] + LANE; e < m.flev_adr[l + 1]; e += WAVE) {
      int i = m.flev_i[e], j = m.flev_j[e];
      float acc = 0.f;
      if (sparse) {  // row k: its diagonal, then its ancestors by decreasing depth
        for (int q = m.flev_kadr[e]; q < m.flev_kadr[e] + m.flev_kcnt[e]; q++) {
          int k = m.flev_k[q], d = m.dof_depth[k];
          const float* rk = LD + m.dof_madr[k];
          float dk = rk[0];
          if (!(dk > KMINVAL)) { bad = 1; dk = KMINVAL; }
          acc += rk[d - m.dof_depth[i]] * rk[d - m.dof_depth[j]] / dk;
        }
        LD[mxi(m, i, j)] -= acc;
      } else {
        for (int q = m.flev_kadr[e]; q < m.flev_kadr[e] + m.flev_kcnt[e]; q++) {
          int k = m.flev_k[q];
          float dk = LD[k * nv + k];
          if (!(dk > KMINVAL)) { bad = 1; dk = KMINVAL; }
          acc += LD[k * nv + i] * LD[k * nv + j] / dk;
        }
        LD[i * nv + j] -= acc;
      }
    }
    KSYNC();
  }
  if (sparse) {
    for (int k = LANE; k < nv; k += WAVE) {
      float* row = LD + m.dof_madr[k];  // diagonal, then the ancestors
      float dk = row[0];
      if (!(dk > KMINVAL)) { bad = 1; dk = KMINVAL; }
      for (int a = 1, n = m.dof_nchain[k]; a < n; a++) row[a] /= dk;
    }
  } else {
    for (int t = LANE; t < nv * nv; t += WAVE) {
      int k = t / nv, i = t - k * nv;
      if (i < k) {
        float dk = LD[k * nv + k];
        LD[t] = dk > KMINVAL ? LD[t] / dk : LD[t] / KMINVAL;
      } else if (i == k && !(LD[t] > KMINVAL)) {
        bad = 1;
      }
    }
  }
  KSYNC();
  return wori(bad) ? ST_M_NOT_PD : 0;
}

// Path trees (DModel::ntree > 0, compact rows): dof j of a tree starting at
// dof d0 has the ancestors j - 1, .., 0 of the same tree, so its compact row
// (diagonal, then the ancestors nearest first) is entry (j, i) at
// j (j + 1) / 2 + j - i of the tree's block, which starts at dof_madr[d0].
KDEV bool path_trees(const DModel& m) { return m_compact(m) && m.ntree > 0; }
KDEV int ptx(int j, int i) { return j * (j + 1) / 2 + j - i; }

// x <- (L^T D L)^-1 x on path trees: one lane per tree, the tree's entries of
// x in registers. The same operations in the same order as the level-parallel
// solve below (each dof's sum over its descendants, then over its ancestors
// nearest first), so the results are the same bits.
KDEV void solve_ld_trees(const DModel& m, const float* LD, float* x) {
  for (int t = LANE; t < m.ntree; t += WAVE) {
    int d0 = m.tree_adr[t], n = m.tree_adr[t + 1] - d0;
    const float* T = LD + m.dof_madr[d0];
    float v[PATH_TREE_MAX];
#pragma unroll
    for (int j = 0; j < PATH_TREE_MAX; j++) v[j] = j < n ? x[d0 + j] : 0.f;
#pragma unroll
    for (int i = PATH_TREE_MAX - 2; i >= 0; i--) {  // L^T y = x, deepest first
      if (i + 1 < n) {
        float acc = 0.f;
#pragma unroll
        for (int k = i + 1; k < PATH_TREE_MAX; k++)
          if (k < n) acc += T[ptx(k, i)] * v[k];
        v[i] -= acc;
      }
    }
#pragma unroll
    for (int j = 0; j < PATH_TREE_MAX; j++)
      if (j < n) v[j] /= T[ptx(j, j)];
#pragma unroll
    for (int k = 1; k < PATH_TREE_MAX; k++) {  // L z = y, from the root
      if (k < n) {
        float acc = 0.f;
#pragma unroll
        for (int i = k - 1; i >= 0; i--) acc += T[ptx(k, i)] * v[i];
        v[k] -= acc;
      }
    }
#pragma unroll
    for (int j = 0; j < PATH_TREE_MAX; j++)
      if (j < n) x[d0 + j] = v[j];
  }
  KSYNC();
}

// y = M x on path trees, one lane per tree; per dof the order of mulM_row
// (ancestors nearest first, then the diagonal and the descendants)
KDEV void mulM_trees(const DModel& m, const float* M, const float* x, float* y) {
  for (int t = LANE; t < m.ntree; t += WAVE) {
    int d0 = m.tree_adr[t], n = m.tree_adr[t + 1] - d0;
    const float* T = M + m.dof_madr[d0];
    float v[PATH_TREE_MAX];
#pragma unroll
    for (int j = 0; j < PATH_TREE_MAX; j++) v[j] = j < n ? x[d0 + j] : 0.f;
#pragma unroll
    for (int i = 0; i < PATH_TREE_MAX; i++) {
      if (i < n) {
        float acc = 0.f;
#pragma unroll
        for (int j = i - 1; j >= 0; j--) acc += T[ptx(i, j)] * v[j];
#pragma unroll
        for (int k = i; k < PATH_TREE_MAX; k++)
          if (k < n) acc += T[ptx(k, i)] * v[k];
        y[d0 + i] = acc;
      }
    }
  }
}

// x <- (L^T D L)^-1 x with L unit lower (entries outside the dof tree are 0).
// GPU, nv <= 64: lane i holds x_i; each column step broadcasts one solved entry
// with readlane and updates the other lanes, with no LDS round trip per step.
// Otherwise one workspace phase per tree level: L^T y = x deepest level first,
// each dof gathering from its (final) descendants, a contiguous dof range;
// L z = y from the roots, each dof gathering from its (final) ancestors.
KDEV void solve_ld(const DModel& m, const float* LD, float* x) {
  int nv = m.nv;
#if WAVE >= 32
  if (!MJK_BIG || nv <= WAVE) {
    // the next step's factor entry is loaded ahead of the current step's
    // readlane -> FMA chain, so the LDS latency overlaps it
    int ln = LANE < nv ? LANE : 0;
    float v = LANE < nv ? x[LANE] : 0.f;
    float ln_next = nv > 1 ? LD[(nv - 1) * nv + ln] : 0.f;
    for (int k = nv - 1; k >= 1; k--) {  // L^T y = x: column k updates rows i < k
      float l = ln_next;
      if (k > 1) ln_next = LD[(k - 1) * nv + ln];
      float xk = rl_f(v, k);
      if (LANE < k) v -= l * xk;
    }
    if (LANE < nv) v /= LD[LANE * nv + LANE];
    const float* LDr = LD + ln * nv;
    ln_next = LDr[0];
    for (int i = 0; i < nv - 1; i++) {   // L z = y: solved entry i updates rows k > i
      float l = ln_next;
      if (i + 1 < nv - 1) ln_next = LDr[i + 1];
      float xi = rl_f(v, i);
      if (LANE > i && LANE < nv) v -= l * xi;
    }
    KSYNC();
    if (LANE < nv) x[LANE] = v;
    KSYNC();
    return;
  }
#endif
  if (path_trees(m)) {
    solve_ld_trees(m, LD, x);
    return;
  }
  for (int lv = m.ndoflevel - 2; lv >= 0; lv--) {
    for (int t = m.doflevel_adr[lv] + LANE; t < m.doflevel_adr[lv + 1]; t += WAVE) {
      int i = m.doflevel_dof[t];
      float acc = 0.f;
      for (int k = i + 1; k < m.dof_subtree_end[i]; k++) acc += LD[mxi(m, k, i)] * x[k];
      x[i] -= acc;
    }
    KSYNC();
  }
  for (int i = LANE; i < nv; i += WAVE) x[i] /= LD[mxi(m, i, i)];
  KSYNC();
  for (int lv = 1; lv < m.ndoflevel; lv++) {
    for (int t = m.doflevel_adr[lv] + LANE; t < m.doflevel_adr[lv + 1]; t += WAVE) {
      int k = m.doflevel_dof[t];
      float acc = 0.f;
      for (int i = m.dof_parentid[k]; i >= 0; i = m.dof_parentid[i]) acc += LD[mxi(m, k, i)] * x[i];
      x[k] -= acc;
    }
    KSYNC();
  }
}

// mj_comVel with MuJoCo's grouping of free / ball dofs (reference forward.hpp:287-332)
KDEV void st_com_vel(const DModel& m, const Layout& L, float* w) {
  float *cvel = w + L.cvel, *cdd = w + L.cdofdot;
  const float *cdof = w + L.cdof, *qvel = w + L.qvel;
  if (LANE == 0) for (int k = 0; k < 6; k++) cvel[k] = 0.f;
  KSYNC();
  for (int lv = 0; lv < m.nlevel; lv++) {
    int a = m.level_adr[lv], n = m.level_adr[lv + 1] - a;
    for (int i = LANE; i < n; i += WAVE) {
      int b = m.level_body[a + i];
      float v[6];
      for (int k = 0; k < 6; k++) v[k] = cvel[6 * m.body_parentid[b] + k];
      int ja = m.body_jntadr[b], jn = m.body_jntnum[b];
      for (int j = ja; j < ja + jn; j++) {
        int da = m.jnt_dofadr[j], jt = m.jnt_type[j];
        if (jt == JFREE) {
          for (int q = 0; q < 3; q++) for (int c = 0; c < 6; c++) cdd[6 * (da + q) + c] = 0.f;
          for (int q = 0; q < 3; q++) for (int c = 0; c < 6; c++) v[c] += cdof[6 * (da + q) + c] * qvel[da + q];
          da += 3;
        }
        if (jt == JFREE || jt == JBALL) {
          for (int q = 0; q < 3; q++) cross_motion(v, cdof + 6 * (da + q), cdd + 6 * (da + q));
          for (int q = 0; q < 3; q++) for (int c = 0; c < 6; c++) v[c] += cdof[6 * (da + q) + c] * qvel[da + q];
        } else {
          cross_motion(v, cdof + 6 * da, cdd + 6 * da);
          for (int c = 0; c < 6; c++) v[c] += cdof[6 * da + c] * qvel[da];
        }
      }
      for (int k = 0; k < 6; k++) cvel[6 * b + k] = v[k];
    }
    KSYNC();
  }
}

// mj_rne (zero acceleration) -> qfrc_bias, then passive and actuator forces,
// qfrc_smooth and qacc_smooth = M^-1 qfrc_smooth.
// Reference forward.hpp:338-469.
// The env's friction scale: under domain randomisation all five friction
// coefficients of every pair are multiplied by it, otherwise it is exactly 1
// (reference src/compat/mkx_kernels/solver.hpp:146,435-436,530). Callers
// read a pair's coefficients from the model (m.pair_friction + 5 p) and scale
// each value they use, so the reads stay global-memory loads: a pointer that
// could also point at a private copy made them flat loads with the copy in
// scratch memory, in the elliptic cone's cost and line search.
KDEV float fric_scale(const DModel& m, const Layout& L, const float* w) {
  return L.rand >= 0 ? w[L.rand + rand_fric(m.nbody)] : 1.f;
}

// Quadratic drag of a free joint's translation under domain randomisation:
// F = -k |v| v on its three translational dofs, k per env and free joint
// (reference src/compat/mkx_kernels/forward.hpp:455-465). Returns the
// force on dof k (0 for other dofs).
KDEV float free_drag(const DModel& m, const Layout& L, const float* w, int k) {
  int j = m.dof_jntid[k], i = k - m.jnt_dofadr[j];
  if (m.jnt_type[j] != JFREE || i >= 3) return 0.f;
  int f = 0;
  for (int t = 0; t < j; t++) f += m.jnt_type[t] == JFREE;
  const float* v = w + L.qvel + m.jnt_dofadr[j];
  float kd = -w[L.rand + rand_drag(m.nbody, m.nu) + f] * sqrtf(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
  return kd * v[i];
}

KDEV int st_smooth(const DModel& m, const Layout& L, float* w) {
  float *cacc = w + L.cacc, *cfrc = w + L.cfrc;
  const float *cinert = w + L.cinert, *cvel = w + L.cvel, *cdd = w + L.cdofdot, *cdof = w + L.cdof;
  const float *qvel = w + L.qvel, *qpos = w + L.qpos, *ctrl = w + L.ctrl;
  if (LANE == 0) {
    for (int k = 0; k < 6; k++) cacc[k] = 0.f;
    if (!(m.disableflags & DS_GRAVITY)) { cacc[3] = -m.gravity[0]; cacc[4] = -m.gravity[1]; cacc[5] = -m.gravity[2]; }
    for (int k = 0; k < 6; k++) cfrc[k] = 0.f;
  }
  KSYNC();
  for (int lv = 0; lv < m.nlevel; lv++) {
    int a = m.level_adr[lv], n = m.level_adr[lv + 1] - a;
    for (int i = LANE; i < n; i += WAVE) {
      int b = m.level_body[a + i];
      float ac[6];
      for (int k = 0; k < 6; k++) ac[k] = cacc[6 * m.body_parentid[b] + k];
      int da = m.body_dofadr[b], dn = m.body_dofnum[b];
      for (int q = da; q < da + dn; q++) for (int c = 0; c < 6; c++) ac[c] += cdd[6 * q + c] * qvel[q];
      for (int k = 0; k < 6; k++) cacc[6 * b + k] = ac[k];
      float f1[6], mv[6], f2[6];
      mul_inert(cinert + 10 * b, ac, f1);
      mul_inert(cinert + 10 * b, cvel + 6 * b, mv);
      cross_force(cvel + 6 * b, mv, f2);
      for (int k = 0; k < 6; k++) cfrc[6 * b + k] = f1[k] + f2[k];
    }
    KSYNC();
  }
  // subtree sums of body forces into cacc (no longer needed), one lane per
  // (body, component)
  for (int t = LANE; t < 6 * m.nbody; t += WAVE) {
    int b = t / 6, c = t - 6 * b;
    float s = 0.f;
    for (int k = b; k < m.body_subtree_end[b]; k++) s += cfrc[6 * k + c];
    cacc[t] = s;
  }
  KSYNC();
  float *qfs = w + L.qfrc_smooth, *qas = w + L.qacc_smooth, *af = w + L.act_force;
  for (int k = LANE; k < m.nv; k += WAVE) {
    const float* cd = cdof + 6 * k;
    const float* f = cacc + 6 * m.dof_bodyid[k];
    float bias = cd[0] * f[0] + cd[1] * f[1] + cd[2] * f[2] + cd[3] * f[3] + cd[4] * f[4] + cd[5] * f[5];
    float pas = 0.f;
    if (!(m.disableflags & DS_PASSIVE)) {
      int j = m.dof_jntid[k], jt = m.jnt_type[j];
      if ((jt == JHINGE || jt == JSLIDE) && m.jnt_stiffness[j] != 0.f) {
        int qa = m.jnt_qposadr[j];
        pas -= m.jnt_stiffness[j] * (qpos[qa] - m.qpos_spring[qa]);
      }
      pas -= m.dof_damping[k] * qvel[k];
    }
    float act = 0.f;
    for (int t = m.dof_act_adr[k]; t < m.dof_act_adr[k] + m.dof_act_num[k]; t++) {
      int u = m.dof_act_list[t];
      float frc = 0.f;
      if (!(m.disableflags & DS_ACTUATION)) {
        float gear = m.act_gear[u];
        float c = ctrl[u];
        if (m.act_ctrllimited[u] && !(m.disableflags & DS_CLAMPCTRL)) c = clampf(c, m.act_ctrlrange[2 * u], m.act_ctrlrange[2 * u + 1]);
        float len = gear * qpos[m.act_qadr[u]], vel = gear * qvel[k];
        const float *g = m.act_gain + 3 * u, *bi = m.act_bias + 3 * u;
        frc = g[0] * c + bi[0] + bi[1] * len + bi[2] * vel;
        if (L.rand >= 0) frc *= w[L.rand + rand_gain(m.nbody) + u];  // randomised actuator strength
        if (m.act_forcelimited[u]) frc = clampf(frc, m.act_forcerange[2 * u], m.act_forcerange[2 * u + 1]);
        act += gear * frc;
      }
      af[u] = frc;
    }
    float s = pas - bias + act;
    if (L.rand >= 0) s += free_drag(m, L, w, k);
    qfs[k] = s;
    qas[k] = s;
  }
  KSYNC();
  solve_ld(m, w + L.qLD, qas);
  return 0;
}

// ------------------------------------------------------------------ collision
struct RawC { float dist, pos[3], n[3]; };

KDEV int k_sphere_sphere(const float* c1, float r1, const float* c2, float r2, float margin, RawC* o) {
  float dv[3] = {c2[0] - c1[0], c2[1] - c1[1], c2[2] - c1[2]};
  float Ls = sqrtf(dot3(dv, dv));
  float dist = Ls - r1 - r2;
  if (dist > margin) return 0;
  float n[3];
  if (Ls < KMINVAL) { n[0] = 1.f; n[1] = 0.f; n[2] = 0.f; }
  else { float in = 1.0f / Ls; n[0] = dv[0] * in; n[1] = dv[1] * in; n[2] = dv[2] * in; }
  o->dist = dist;
  for (int i = 0; i < 3; i++) { o->n[i] = n[i]; o->pos[i] = c1[i] + n[i] * (r1 + 0.5f * dist); }
  return 1;
}
KDEV int k_plane_point(const float* pp, const float* pn, const float* c, float r, float margin, RawC* o) {
  float dv[3] = {c[0] - pp[0], c[1] - pp[1], c[2] - pp[2]};
  float dist = dot3(pn, dv) - r;
  if (dist > margin) return 0;
  o->dist = dist;
  for (int i = 0; i < 3; i++) { o->n[i] = pn[i]; o->pos[i] = c[i] - pn[i] * (r + 0.5f * dist); }
  return 1;
}
KDEV void k_seg_seg(const float* p1, const float* d1, const float* p2, const float* d2, float* so, float* to) {
  float r[3] = {p1[0] - p2[0], p1[1] - p2[1], p1[2] - p2[2]};
  float a = dot3(d1, d1), e = dot3(d2, d2), f = dot3(d2, r);
  float s, t;
  if (a <= KMINVAL && e <= KMINVAL) { s = t = 0.f; }
  else if (a <= KMINVAL) { s = 0.f; t = clampf(f / e, 0.f, 1.f); }
  else {
    float c = dot3(d1, r);
    if (e <= KMINVAL) { t = 0.f; s = clampf(-c / a, 0.f, 1.f); }
    else {
      float b = dot3(d1, d2);
      float denom = a * e - b * b;
      s = (denom > 1e-10f * a * e) ? clampf((b * f - c * e) / denom, 0.f, 1.f) : 0.f;
      t = (b * s + f) / e;
      if (t < 0.f) { t = 0.f; s = clampf(-c / a, 0.f, 1.f); }
      else if (t > 1.f) { t = 1.f; s = clampf((b - c) / a, 0.f, 1.f); }
    }
  }
  *so = s; *to = t;
}
KDEV void k_capsule_ends(const float* x, const float* R, float h, float* p0, float* d) {
  float ax[3] = {R[2], R[5], R[8]};
  for (int i = 0; i < 3; i++) { p0[i] = x[i] - ax[i] * h; d[i] = 2.f * h * ax[i]; }
}
KDEV int k_sphere_box(const float* c, float r, const float* bx, const float* R, const float* s, float margin, RawC* o) {
  float rel[3] = {c[0] - bx[0], c[1] - bx[1], c[2] - bx[2]}, pl[3];
  mtv3(R, rel, pl);
  float q[3];
  bool inside = true;
  for (int i = 0; i < 3; i++) {
    q[i] = clampf(pl[i], -s[i], s[i]);
    if (pl[i] < -s[i] || pl[i] > s[i]) inside = false;
  }
  float nl[3], dist;
  if (!inside) {
    float dv[3] = {pl[0] - q[0], pl[1] - q[1], pl[2] - q[2]};
    float Ls = sqrtf(dot3(dv, dv));
    dist = Ls - r;
    if (dist > margin) return 0;
    for (int i = 0; i < 3; i++) nl[i] = -dv[i] / Ls;
  } else {
    int k = 0;
    float best = s[0] - fabsf(pl[0]);
    for (int i = 1; i < 3; i++) { float dp = s[i] - fabsf(pl[i]); if (dp < best) { best = dp; k = i; } }
    dist = -best - r;
    if (dist > margin) return 0;
    nl[0] = nl[1] = nl[2] = 0.f;
    nl[k] = pl[k] >= 0.f ? -1.f : 1.f;
  }
  float n[3];
  mv3(R, nl, n);
  o->dist = dist;
  for (int i = 0; i < 3; i++) { o->n[i] = n[i]; o->pos[i] = c[i] + n[i] * (r + 0.5f * dist); }
  return 1;
}

// signed distance of a point (box frame) to a box of half-sizes h (negative inside)
KDEV float k_box_sdf(const float* p, const float* h) {
  float q[3] = {fabsf(p[0]) - h[0], fabsf(p[1]) - h[1], fabsf(p[2]) - h[2]};
  float o[3] = {fmaxf(q[0], 0.f), fmaxf(q[1], 0.f), fmaxf(q[2], 0.f)};
  return sqrtf(dot3(o, o)) + fminf(fmaxf(q[0], fmaxf(q[1], q[2])), 0.f);
}

// Capsule (geom1) against box (geom2), up to 2 contacts. The box signed
// distance along the capsule axis is convex, so a golden-section search finds
// its minimum t*; the capsule is then a sphere at t* (and, when they are
// within the margin and apart from t*, at the segment ends). When both ends
// touch and the interior is no deeper, the two ends are the contacts (a
// capsule lying on a face).
KDEV int k_capsule_box(const float* xc, const float* Rc, const float* sc, const float* xb, const float* Rb,
                       const float* sb, float margin, RawC* out) {
  float p0[3], d[3];
  k_capsule_ends(xc, Rc, sc[1], p0, d);
  float r = sc[0];
  float l0[3], ld[3], rel[3] = {p0[0] - xb[0], p0[1] - xb[1], p0[2] - xb[2]};
  mtv3(Rb, rel, l0);
  mtv3(Rb, d, ld);
  const float gr = 0.381966011f;  // 2 - golden ratio
  float a = 0.f, b = 1.f;
  float t1 = a + gr * (b - a), t2 = b - gr * (b - a);
  float q1[3] = {l0[0] + t1 * ld[0], l0[1] + t1 * ld[1], l0[2] + t1 * ld[2]};
  float q2[3] = {l0[0] + t2 * ld[0], l0[1] + t2 * ld[1], l0[2] + t2 * ld[2]};
  float f1 = k_box_sdf(q1, sb), f2 = k_box_sdf(q2, sb);
  for (int it = 0; it < 32; it++) {
    if (f1 <= f2) {
      b = t2; t2 = t1; f2 = f1;
      t1 = a + gr * (b - a);
      float q[3] = {l0[0] + t1 * ld[0], l0[1] + t1 * ld[1], l0[2] + t1 * ld[2]};
      f1 = k_box_sdf(q, sb);
    } else {
      a = t1; t1 = t2; f1 = f2;
      t2 = b - gr * (b - a);
      float q[3] = {l0[0] + t2 * ld[0], l0[1] + t2 * ld[1], l0[2] + t2 * ld[2]};
      f2 = k_box_sdf(q, sb);
    }
  }
  float ts = 0.5f * (a + b);
  float qs[3] = {l0[0] + ts * ld[0], l0[1] + ts * ld[1], l0[2] + ts * ld[2]};
  float qa[3] = {l0[0], l0[1], l0[2]}, qb[3] = {l0[0] + ld[0], l0[1] + ld[1], l0[2] + ld[2]};
  float fs = k_box_sdf(qs, sb), fa = k_box_sdf(qa, sb), fb = k_box_sdf(qb, sb);
  if (fs - r > margin) return 0;
  float tc[2];
  int nc = 0;
  if (fa - r <= margin && fb - r <= margin && fs >= fminf(fa, fb) - 1e-6f) {
    tc[0] = 0.f; tc[1] = 1.f; nc = 2;
  } else {
    tc[nc++] = ts;
    float te = fa <= fb ? 0.f : 1.f, fe = fminf(fa, fb);
    if (fe - r <= margin && fabsf(te - ts) > 0.1f) tc[nc++] = te;
  }
  int n = 0;
  for (int k = 0; k < nc; k++) {
    float c[3] = {p0[0] + tc[k] * d[0], p0[1] + tc[k] * d[1], p0[2] + tc[k] * d[2]};
    n += k_sphere_box(c, r, xb, Rb, sb, margin, out + n);
  }
  return n;
}

// Box (geom1) against box (geom2) by the separating-axis test over the 15
// axes, up to 4 contacts. Face axis: the incident face of the other box is
// clipped against the side planes of the reference face and the points
// below it (within the margin) are kept, reduced to the deepest point plus
// the three that span the largest area. Edge axis: one contact at the
// closest points of the two supporting edges. Face axes are preferred unless
// an edge axis separates clearly more.
KDEV int k_box_box(const float* x1, const float* R1, const float* h1, const float* x2, const float* R2,
                   const float* h2, float margin, RawC* out) {
  float C[9], t[3], dw[3] = {x2[0] - x1[0], x2[1] - x1[1], x2[2] - x1[2]};
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++) C[3 * i + j] = R1[i] * R2[j] + R1[3 + i] * R2[3 + j] + R1[6 + i] * R2[6 + j];
  mtv3(R1, dw, t);  // box 2 centre in box 1's frame; C[i][j] = a_i . b_j
  float best_face = -3.0e38f, best_edge = -3.0e38f;
  int face = -1, edge = -1;
  float edge_ax[3] = {0.f, 0.f, 0.f};
  for (int i = 0; i < 3; i++) {  // axes of box 1
    float rb = h2[0] * fabsf(C[3 * i]) + h2[1] * fabsf(C[3 * i + 1]) + h2[2] * fabsf(C[3 * i + 2]);
    float sep = fabsf(t[i]) - h1[i] - rb;
    if (sep > margin) return 0;
    if (sep > best_face) { best_face = sep; face = i; }
  }
  for (int j = 0; j < 3; j++) {  // axes of box 2
    float ra = h1[0] * fabsf(C[j]) + h1[1] * fabsf(C[3 + j]) + h1[2] * fabsf(C[6 + j]);
    float tb = t[0] * C[j] + t[1] * C[3 + j] + t[2] * C[6 + j];
    float sep = fabsf(tb) - ra - h2[j];
    if (sep > margin) return 0;
    if (sep > best_face) { best_face = sep; face = 3 + j; }
  }
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++) {  // a_i x b_j in box 1's frame
      float bj[3] = {C[j], C[3 + j], C[6 + j]}, ai[3] = {0.f, 0.f, 0.f}, L[3];
      ai[i] = 1.f;
      cross3(ai, bj, L);
      float ln = sqrtf(dot3(L, L));
      if (ln < 1e-5f) continue;  // parallel edges: covered by the face axes
      for (int k = 0; k < 3; k++) L[k] /= ln;
      float ra = h1[0] * fabsf(L[0]) + h1[1] * fabsf(L[1]) + h1[2] * fabsf(L[2]);
      float lb[3] = {L[0] * C[0] + L[1] * C[3] + L[2] * C[6], L[0] * C[1] + L[1] * C[4] + L[2] * C[7],
                     L[0] * C[2] + L[1] * C[5] + L[2] * C[8]};
      float rb = h2[0] * fabsf(lb[0]) + h2[1] * fabsf(lb[1]) + h2[2] * fabsf(lb[2]);
      float sep = fabsf(dot3(t, L)) - ra - rb;
      if (sep > margin) return 0;
      if (sep > best_edge) {
        best_edge = sep; edge = 3 * i + j;
        float sg = dot3(t, L) < 0.f ? -1.f : 1.f;
        edge_ax[0] = sg * L[0]; edge_ax[1] = sg * L[1]; edge_ax[2] = sg * L[2];
      }
    }
  float hmin = fminf(fminf(fminf(h1[0], h1[1]), fminf(h1[2], h2[0])), fminf(h2[1], h2[2]));
  if (edge >= 0 && best_edge > 0.95f * best_face + 0.01f * hmin) {
    // edge-edge: supporting edge of box 1 along +axis, of box 2 along -axis (box 1 frame)
    int i = edge / 3, j = edge % 3;
    float p1[3], p2[3], e1[3] = {0.f, 0.f, 0.f}, e2[3], bj[3] = {C[j], C[3 + j], C[6 + j]};
    for (int k = 0; k < 3; k++) p1[k] = k == i ? 0.f : (edge_ax[k] >= 0.f ? h1[k] : -h1[k]);
    p1[i] = -h1[i];
    e1[i] = 2.f * h1[i];
    float c2[3] = {t[0], t[1], t[2]};
    for (int k = 0; k < 3; k++) {
      if (k == j) continue;
      float bk[3] = {C[k], C[3 + k], C[6 + k]};
      float sg = dot3(bk, edge_ax) > 0.f ? -1.f : 1.f;
      for (int q = 0; q < 3; q++) c2[q] += sg * h2[k] * bk[q];
    }
    for (int q = 0; q < 3; q++) { p2[q] = c2[q] - h2[j] * bj[q]; e2[q] = 2.f * h2[j] * bj[q]; }
    float s, u;
    k_seg_seg(p1, e1, p2, e2, &s, &u);
    float m1[3], mid[3], nw[3];
    for (int q = 0; q < 3; q++) m1[q] = 0.5f * (p1[q] + s * e1[q] + p2[q] + u * e2[q]);
    mv3(R1, m1, mid);
    mv3(R1, edge_ax, nw);
    out->dist = best_edge;
    for (int q = 0; q < 3; q++) { out->pos[q] = mid[q] + x1[q]; out->n[q] = nw[q]; }
    return 1;
  }
  // face contact: reference box owns the axis; normal nr (world) points from it to the incident box
  bool ref1 = face < 3;
  int ax = ref1 ? face : face - 3;
  const float *xr = ref1 ? x1 : x2, *Rr = ref1 ? R1 : R2, *hr = ref1 ? h1 : h2;
  const float *xi = ref1 ? x2 : x1, *Ri = ref1 ? R2 : R1, *hi = ref1 ? h2 : h1;
  float dri[3] = {xi[0] - xr[0], xi[1] - xr[1], xi[2] - xr[2]};
  float nr[3] = {Rr[ax], Rr[3 + ax], Rr[6 + ax]};
  if (dot3(nr, dri) < 0.f) { nr[0] = -nr[0]; nr[1] = -nr[1]; nr[2] = -nr[2]; }
  // incident face: the face of the other box most anti-parallel to nr
  int fi = 0;
  float fmin = 3.0e38f, fsg = 1.f;
  for (int k = 0; k < 3; k++) {
    float dk = Ri[k] * nr[0] + Ri[3 + k] * nr[1] + Ri[6 + k] * nr[2];
    if (-fabsf(dk) < fmin) { fmin = -fabsf(dk); fi = k; fsg = dk > 0.f ? -1.f : 1.f; }
  }
  // incident face corners in the reference box frame
  float poly[8][3], tmp[8][3];
  int np = 4;
  {
    int u = (fi + 1) % 3, v = (fi + 2) % 3;
    const float su[4] = {-1.f, 1.f, 1.f, -1.f}, sv[4] = {-1.f, -1.f, 1.f, 1.f};
    for (int c = 0; c < 4; c++) {
      float pl[3] = {0.f, 0.f, 0.f}, pw[3], rl[3];
      pl[fi] = fsg * hi[fi]; pl[u] = su[c] * hi[u]; pl[v] = sv[c] * hi[v];
      mv3(Ri, pl, pw);
      for (int q = 0; q < 3; q++) pw[q] += xi[q] - xr[q];
      mtv3(Rr, pw, rl);
      for (int q = 0; q < 3; q++) poly[c][q] = rl[q];
    }
  }
  // clip against the four side planes |p[k]| <= hr[k], k != ax
  for (int k = 0; k < 3; k++) {
    if (k == ax) continue;
    for (int sg = -1; sg <= 1; sg += 2) {
      int nn = 0;
      for (int c = 0; c < np; c++) {
        const float* A = poly[c];
        const float* Bp = poly[(c + 1) % np];
        float da = sg * A[k] - hr[k], db = sg * Bp[k] - hr[k];
        if (da <= 0.f && nn < 8) { for (int q = 0; q < 3; q++) tmp[nn][q] = A[q]; nn++; }
        if ((da < 0.f) != (db < 0.f) && nn < 8) {
          float tt = da / (da - db);
          for (int q = 0; q < 3; q++) tmp[nn][q] = A[q] + tt * (Bp[q] - A[q]);
          nn++;
        }
      }
      np = nn;
      for (int c = 0; c < np; c++) for (int q = 0; q < 3; q++) poly[c][q] = tmp[c][q];
    }
  }
  // depth of each clipped point below the reference face (normal along +/- axis ax)
  float sgn = 0.f;
  {
    float nl[3];
    mtv3(Rr, nr, nl);
    sgn = nl[ax] >= 0.f ? 1.f : -1.f;
  }
  float dep[8];
  int keep[8], nk = 0;
  for (int c = 0; c < np; c++) {
    dep[c] = sgn * poly[c][ax] - hr[ax];
    if (dep[c] <= margin) keep[nk++] = c;
  }
  if (nk == 0) return 0;
  int sel[4], ns = 0;
  if (nk <= 4) {
    for (int c = 0; c < nk; c++) sel[ns++] = keep[c];
  } else {
    // deepest, farthest from it, then the largest triangle on each side of that chord
    int i0 = keep[0];
    for (int c = 1; c < nk; c++) if (dep[keep[c]] < dep[i0]) i0 = keep[c];
    int i1 = -1;
    float dbest = -1.f;
    for (int c = 0; c < nk; c++) {
      float dd[3] = {poly[keep[c]][0] - poly[i0][0], poly[keep[c]][1] - poly[i0][1], poly[keep[c]][2] - poly[i0][2]};
      float l2 = dot3(dd, dd);
      if (l2 > dbest) { dbest = l2; i1 = keep[c]; }
    }
    float e[3] = {poly[i1][0] - poly[i0][0], poly[i1][1] - poly[i0][1], poly[i1][2] - poly[i0][2]};
    int i2 = -1, i3 = -1;
    float amax = 0.f, amin = 0.f;
    for (int c = 0; c < nk; c++) {
      float f[3] = {poly[keep[c]][0] - poly[i0][0], poly[keep[c]][1] - poly[i0][1], poly[keep[c]][2] - poly[i0][2]}, cr[3];
      cross3(e, f, cr);
      float area = cr[ax] * sgn;
      if (area > amax) { amax = area; i2 = keep[c]; }
      if (area < amin) { amin = area; i3 = keep[c]; }
    }
    sel[ns++] = i0; sel[ns++] = i1;
    if (i2 >= 0) sel[ns++] = i2;
    if (i3 >= 0) sel[ns++] = i3;
  }
  float nw[3] = {ref1 ? nr[0] : -nr[0], ref1 ? nr[1] : -nr[1], ref1 ? nr[2] : -nr[2]};
  for (int c = 0; c < ns; c++) {
    const float* pl = poly[sel[c]];
    float d = dep[sel[c]];
    float mid[3] = {pl[0], pl[1], pl[2]}, mw[3];
    mid[ax] -= 0.5f * sgn * d;  // halfway between the point and the reference face
    mv3(Rr, mid, mw);
    out[c].dist = d;
    for (int q = 0; q < 3; q++) { out[c].pos[q] = mw[q] + xr[q]; out[c].n[q] = nw[q]; }
  }
  return ns;
}

// ------------------------------------------------------------------ GJK / EPA
// Convex primitives (sphere, capsule, ellipsoid, cylinder, box) through their
// support functions: GJK (Gilbert, Johnson and Keerthi 1988; Ericson's
// Voronoi-region nearest points on the simplex) gives the distance of
// separated shapes; for overlapping ones EPA (van den Bergen 2001) grows the
// final simplex into a polytope of the Minkowski difference A - B until its
// face nearest the origin lies on the boundary. One contact; normal from A
// (geom1) to B (geom2). The polytope lives in per-lane private memory.
struct KCvx { int type; const float *x, *R, *s; };
struct KSv { float w[3], a[3], b[3]; };

KDEV void k_support(const KCvx& g, const float* d, float* out) {
  float dl[3], pl[3] = {0.f, 0.f, 0.f};
  mtv3(g.R, d, dl);
  float n = sqrtf(dot3(dl, dl));
  if (n < 1e-30f) { dl[0] = 1.f; dl[1] = dl[2] = 0.f; n = 1.f; }
  const float* s = g.s;
  if (g.type == GSPHERE || g.type == GCAPSULE) {
    for (int i = 0; i < 3; i++) pl[i] = s[0] * dl[i] / n;
    if (g.type == GCAPSULE) pl[2] += dl[2] >= 0.f ? s[1] : -s[1];
  } else if (g.type == GELLIPSOID) {
    float den = sqrtf(s[0] * s[0] * dl[0] * dl[0] + s[1] * s[1] * dl[1] * dl[1] + s[2] * s[2] * dl[2] * dl[2]);
    for (int i = 0; i < 3; i++) pl[i] = s[i] * s[i] * dl[i] / den;
  } else if (g.type == GCYLINDER) {
    float r = sqrtf(dl[0] * dl[0] + dl[1] * dl[1]);
    if (r > 1e-30f) { pl[0] = s[0] * dl[0] / r; pl[1] = s[0] * dl[1] / r; }
    pl[2] = dl[2] >= 0.f ? s[1] : -s[1];
  } else {
    for (int i = 0; i < 3; i++) pl[i] = dl[i] >= 0.f ? s[i] : -s[i];
  }
  mv3(g.R, pl, out);
  for (int i = 0; i < 3; i++) out[i] += g.x[i];
}

KDEV void k_msupport(const KCvx& A, const KCvx& B, const float* d, KSv* v) {
  float nd[3] = {-d[0], -d[1], -d[2]};
  k_support(A, d, v->a);
  k_support(B, nd, v->b);
  for (int i = 0; i < 3; i++) v->w[i] = v->a[i] - v->b[i];
}

// barycentric coordinates of the point of triangle (a, b, c) nearest the origin
KDEV void k_tri_nearest(const float* a, const float* b, const float* c, float* l) {
  float ab[3], ac[3], ap[3] = {-a[0], -a[1], -a[2]}, bp[3] = {-b[0], -b[1], -b[2]}, cp[3] = {-c[0], -c[1], -c[2]};
  for (int i = 0; i < 3; i++) { ab[i] = b[i] - a[i]; ac[i] = c[i] - a[i]; }
  float d1 = dot3(ab, ap), d2 = dot3(ac, ap);
  if (d1 <= 0.f && d2 <= 0.f) { l[0] = 1.f; l[1] = l[2] = 0.f; return; }
  float d3 = dot3(ab, bp), d4 = dot3(ac, bp);
  if (d3 >= 0.f && d4 <= d3) { l[1] = 1.f; l[0] = l[2] = 0.f; return; }
  float vc = d1 * d4 - d3 * d2;
  if (vc <= 0.f && d1 >= 0.f && d3 <= 0.f) { float v = d1 / (d1 - d3); l[0] = 1.f - v; l[1] = v; l[2] = 0.f; return; }
  float d5 = dot3(ab, cp), d6 = dot3(ac, cp);
  if (d6 >= 0.f && d5 <= d6) { l[2] = 1.f; l[0] = l[1] = 0.f; return; }
  float vb = d5 * d2 - d1 * d6;
  if (vb <= 0.f && d2 >= 0.f && d6 <= 0.f) { float w = d2 / (d2 - d6); l[0] = 1.f - w; l[1] = 0.f; l[2] = w; return; }
  float va = d3 * d6 - d5 * d4;
  if (va <= 0.f && d4 - d3 >= 0.f && d5 - d6 >= 0.f) {
    float w = (d4 - d3) / ((d4 - d3) + (d5 - d6));
    l[0] = 0.f; l[1] = 1.f - w; l[2] = w; return;
  }
  float den = 1.f / (va + vb + vc), v = vb * den, w = vc * den;
  l[0] = 1.f - v - w; l[1] = v; l[2] = w;
}

// nearest point of the simplex to the origin; keeps the vertices with non-zero
// weight; returns 1 when a tetrahedron holds the origin
KDEV int k_simplex_nearest(KSv* S, int* n, float* lam, float* v) {
  float l[4] = {0.f, 0.f, 0.f, 0.f};
  if (*n == 1) {
    l[0] = 1.f;
  } else if (*n == 2) {
    float e[3], ee = 0.f, t = 0.f;
    for (int i = 0; i < 3; i++) { e[i] = S[1].w[i] - S[0].w[i]; ee += e[i] * e[i]; t -= S[0].w[i] * e[i]; }
    t = ee > 0.f ? t / ee : 0.f;
    if (t <= 0.f) l[0] = 1.f;
    else if (t >= 1.f) l[1] = 1.f;
    else { l[0] = 1.f - t; l[1] = t; }
  } else if (*n == 3) {
    k_tri_nearest(S[0].w, S[1].w, S[2].w, l);
  } else {
    const int F[4][4] = {{0, 1, 2, 3}, {0, 1, 3, 2}, {0, 2, 3, 1}, {1, 2, 3, 0}};
    float best = 3.0e38f;
    int outside = 0;
    for (int f = 0; f < 4; f++) {
      const float *a = S[F[f][0]].w, *b = S[F[f][1]].w, *c = S[F[f][2]].w, *o = S[F[f][3]].w;
      float ab[3], ac[3], nrm[3], ao[3];
      for (int i = 0; i < 3; i++) { ab[i] = b[i] - a[i]; ac[i] = c[i] - a[i]; ao[i] = o[i] - a[i]; }
      cross3(ab, ac, nrm);
      float so = -dot3(nrm, a), sl = dot3(nrm, ao);
      if (so * sl >= 0.f) continue;
      outside = 1;
      float tl[3], p[3];
      k_tri_nearest(a, b, c, tl);
      for (int i = 0; i < 3; i++) p[i] = tl[0] * a[i] + tl[1] * b[i] + tl[2] * c[i];
      float d = dot3(p, p);
      if (d < best) {
        best = d;
        l[0] = l[1] = l[2] = l[3] = 0.f;
        l[F[f][0]] = tl[0]; l[F[f][1]] = tl[1]; l[F[f][2]] = tl[2];
      }
    }
    if (!outside) { v[0] = v[1] = v[2] = 0.f; return 1; }
  }
  v[0] = v[1] = v[2] = 0.f;
  int k = 0;
  for (int i = 0; i < *n; i++) {
    if (l[i] <= 0.f) continue;
    for (int q = 0; q < 3; q++) v[q] += l[i] * S[i].w[q];
    S[k] = S[i];
    lam[k++] = l[i];
  }
  *n = k;
  return 0;
}

enum { KEPA_MAXV = 64, KEPA_MAXF = 128 };

KDEV void k_epa_face(const KSv* V, int a, int b, int c, int (*F)[3], float (*Fn)[4], int* nf) {
  float e1[3], e2[3], n[3];
  for (int q = 0; q < 3; q++) { e1[q] = V[b].w[q] - V[a].w[q]; e2[q] = V[c].w[q] - V[a].w[q]; }
  cross3(e1, e2, n);
  float l = sqrtf(dot3(n, n));
  if (l > 1e-30f) for (int q = 0; q < 3; q++) n[q] /= l;
  int f = *nf;
  F[f][0] = a; F[f][1] = b; F[f][2] = c;
  Fn[f][0] = n[0]; Fn[f][1] = n[1]; Fn[f][2] = n[2];
  Fn[f][3] = dot3(n, V[a].w);
  *nf = f + 1;
}

KDEV int k_gjk_epa(const KCvx& A, const KCvx& B, float margin, RawC* out) {
  KSv S[4];
  int n = 1;
  float d0[3] = {B.x[0] - A.x[0], B.x[1] - A.x[1], B.x[2] - A.x[2]};
  if (dot3(d0, d0) < 1e-20f) { d0[0] = 1.f; d0[1] = d0[2] = 0.f; }
  k_msupport(A, B, d0, &S[0]);
  float v[3] = {S[0].w[0], S[0].w[1], S[0].w[2]}, lam[4] = {1.f, 0.f, 0.f, 0.f};
  // FP32 resolves the distance to about 1e-6 of the difference body's extent:
  // the loop stops once the distance bound |v| - v.w/|v| is below 1e-5 of it
  // (or 1e-6 relative), before a near-coplanar vertex can make a tetrahedron
  // that seems to hold the origin
  // An origin within 1e-6 of the extent of the simplex counts as touching.
  // Every point s of the current simplex has v.s >= v.v, so the new simplex can
  // only hold the origin when v.w < 0, and |v| only shrinks: a containment
  // claimed with v.w > 0, or a step that does not shrink |v|, is rounding,
  // and the previous simplex is kept as the separated answer.
  float gscale = fmaxf(sqrtf(dot3(v, v)), 1e-6f);
  float touch = 1e-12f * gscale * gscale;
  int overlap = 0;
  for (int it = 0; it < 64; it++) {
    float vv = dot3(v, v);
    if (vv < touch) { overlap = 1; break; }
    KSv w;
    float nd[3] = {-v[0], -v[1], -v[2]};
    k_msupport(A, B, nd, &w);
    float vw = dot3(v, w.w);
    if (vv - vw <= fmaxf(1e-6f * vv, 1e-5f * gscale * sqrtf(vv))) break;
    bool dup = false;
    for (int i = 0; i < n; i++)
      if (fabsf(S[i].w[0] - w.w[0]) + fabsf(S[i].w[1] - w.w[1]) + fabsf(S[i].w[2] - w.w[2]) < 1e-7f * gscale) dup = true;
    if (dup) break;
    KSv Sp[4];
    float lp[4] = {lam[0], lam[1], lam[2], lam[3]}, vp[3] = {v[0], v[1], v[2]};
    int np = n;
    for (int i = 0; i < n; i++) Sp[i] = S[i];
    S[n++] = w;
    int inside = k_simplex_nearest(S, &n, lam, v);
    if (inside && vw < 0.f) { overlap = 1; break; }
    if (inside || !(dot3(v, v) < vv)) {
      n = np;
      for (int i = 0; i < np; i++) { S[i] = Sp[i]; lam[i] = lp[i]; }
      for (int q = 0; q < 3; q++) v[q] = vp[q];
      break;
    }
  }
  if (!overlap) {
    float pa[3] = {0.f, 0.f, 0.f}, pb[3] = {0.f, 0.f, 0.f};
    for (int i = 0; i < n; i++)
      for (int q = 0; q < 3; q++) { pa[q] += lam[i] * S[i].a[q]; pb[q] += lam[i] * S[i].b[q]; }
    float dv[3] = {pb[0] - pa[0], pb[1] - pa[1], pb[2] - pa[2]};
    float dist = sqrtf(dot3(dv, dv));
    if (dist > margin || dist <= 0.f) return 0;
    out->dist = dist;
    for (int q = 0; q < 3; q++) { out->n[q] = dv[q] / dist; out->pos[q] = 0.5f * (pa[q] + pb[q]); }
    return 1;
  }
  KSv V[KEPA_MAXV];
  int nvtx = n;
  for (int i = 0; i < n; i++) V[i] = S[i];
  const float ax[6][3] = {{1, 0, 0}, {-1, 0, 0}, {0, 1, 0}, {0, -1, 0}, {0, 0, 1}, {0, 0, -1}};
  for (int k = 0; nvtx < 4 && k < 6; k++) {
    float dir[3] = {ax[k][0], ax[k][1], ax[k][2]};
    if (nvtx == 2) {
      float e[3] = {V[1].w[0] - V[0].w[0], V[1].w[1] - V[0].w[1], V[1].w[2] - V[0].w[2]};
      cross3(e, ax[k], dir);
    } else if (nvtx == 3) {
      float e1[3], e2[3];
      for (int q = 0; q < 3; q++) { e1[q] = V[1].w[q] - V[0].w[q]; e2[q] = V[2].w[q] - V[0].w[q]; }
      cross3(e1, e2, dir);
      if (k & 1) for (int q = 0; q < 3; q++) dir[q] = -dir[q];
    }
    if (dot3(dir, dir) < 1e-20f) continue;
    KSv w;
    k_msupport(A, B, dir, &w);
    float ok;
    if (nvtx == 1) {
      float e[3] = {w.w[0] - V[0].w[0], w.w[1] - V[0].w[1], w.w[2] - V[0].w[2]};
      ok = dot3(e, e);
    } else if (nvtx == 2) {
      float e1[3], e2[3], c[3];
      for (int q = 0; q < 3; q++) { e1[q] = V[1].w[q] - V[0].w[q]; e2[q] = w.w[q] - V[0].w[q]; }
      cross3(e1, e2, c);
      ok = dot3(c, c);
    } else {
      float e1[3], e2[3], e3[3], c[3];
      for (int q = 0; q < 3; q++) { e1[q] = V[1].w[q] - V[0].w[q]; e2[q] = V[2].w[q] - V[0].w[q]; e3[q] = w.w[q] - V[0].w[q]; }
      cross3(e1, e2, c);
      ok = fabsf(dot3(c, e3));
    }
    if (ok > 1e-14f) V[nvtx++] = w;
  }
  if (nvtx < 4) {  // degenerate: touching
    float dn = sqrtf(dot3(d0, d0));
    out->dist = 0.f;
    for (int q = 0; q < 3; q++) { out->n[q] = d0[q] / dn; out->pos[q] = 0.5f * (V[0].a[q] + V[0].b[q]); }
    return 1;
  }
  // faces keep an outward winding: the first four face away from the opposite
  // vertex, a new face takes the winding of the removed face it replaces
  int F[KEPA_MAXF][3];
  float Fn[KEPA_MAXF][4];
  int nf = 0;
  const int T[4][4] = {{0, 1, 2, 3}, {0, 3, 1, 2}, {0, 2, 3, 1}, {1, 3, 2, 0}};
  for (int f = 0; f < 4; f++) {
    int a = T[f][0], b = T[f][1], c = T[f][2], o = T[f][3];
    float e1[3], e2[3], nn[3], eo[3];
    for (int q = 0; q < 3; q++) { e1[q] = V[b].w[q] - V[a].w[q]; e2[q] = V[c].w[q] - V[a].w[q]; eo[q] = V[o].w[q] - V[a].w[q]; }
    cross3(e1, e2, nn);
    if (dot3(nn, eo) > 0.f) k_epa_face(V, a, c, b, F, Fn, &nf);
    else k_epa_face(V, a, b, c, F, Fn, &nf);
  }
  // FP32 termination: the support gap is resolved to about 1e-7 of the
  // difference body's extent, so the loop stops at a gap of 1e-6 of it (or
  // 1e-5 of the depth); the nearest-face distance is a lower bound on the depth
  // that only grows in exact arithmetic, so a drop marks rounding trouble and
  // the previous face is kept
  float scale = 0.f;
  for (int i = 0; i < 4; i++) scale = fmaxf(scale, sqrtf(dot3(V[i].w, V[i].w)));
  float gtol = 1e-6f * scale, vtol = 1e-7f * scale;
  int pv[3] = {0, 0, 0};
  float pn[4] = {0.f, 0.f, 0.f, -3.0e38f};
  for (int it = 0; it < 48; it++) {
    int best = 0;
    for (int f = 1; f < nf; f++) if (Fn[f][3] < Fn[best][3]) best = f;
    if (Fn[best][3] < pn[3] - gtol) break;
    for (int q = 0; q < 3; q++) pv[q] = F[best][q];
    for (int q = 0; q < 4; q++) pn[q] = Fn[best][q];
    KSv w;
    k_msupport(A, B, Fn[best], &w);
    const float* wa = V[F[best][0]].w;
    float dw[3] = {w.w[0] - wa[0], w.w[1] - wa[1], w.w[2] - wa[2]};
    float gap = dot3(Fn[best], dw);
    if (gap <= fmaxf(1e-5f * Fn[best][3], gtol) || nvtx >= KEPA_MAXV) break;
    int E[KEPA_MAXF][2], ne = 0, keep = 0;
    bool full = false;
    for (int f = 0; f < nf; f++) {
      float rel[3] = {w.w[0] - V[F[f][0]].w[0], w.w[1] - V[F[f][0]].w[1], w.w[2] - V[F[f][0]].w[2]};
      if (dot3(Fn[f], rel) > vtol) {
        for (int e = 0; e < 3; e++) {
          int a = F[f][e], b = F[f][(e + 1) % 3], found = -1;
          for (int q = 0; q < ne; q++) if (E[q][0] == b && E[q][1] == a) { found = q; break; }
          if (found >= 0) { E[found][0] = E[ne - 1][0]; E[found][1] = E[ne - 1][1]; ne--; }
          else if (ne < KEPA_MAXF) { E[ne][0] = a; E[ne][1] = b; ne++; }
          else full = true;
        }
      } else {
        if (keep != f) {
          for (int q = 0; q < 3; q++) F[keep][q] = F[f][q];
          for (int q = 0; q < 4; q++) Fn[keep][q] = Fn[f][q];
        }
        keep++;
      }
    }
    nf = keep;
    if (full || nf + ne > KEPA_MAXF) break;
    V[nvtx] = w;
    for (int q = 0; q < ne; q++) k_epa_face(V, E[q][0], E[q][1], nvtx, F, Fn, &nf);
    nvtx++;
  }
  // the face used: the last accepted nearest face (pv, pn)
  float depth = pn[3];
  float p[3] = {pn[0] * depth, pn[1] * depth, pn[2] * depth};
  const float *a = V[pv[0]].w, *b = V[pv[1]].w, *c = V[pv[2]].w;
  float v0[3], v1[3], v2[3];
  for (int q = 0; q < 3; q++) { v0[q] = b[q] - a[q]; v1[q] = c[q] - a[q]; v2[q] = p[q] - a[q]; }
  float d00 = dot3(v0, v0), d01 = dot3(v0, v1), d11 = dot3(v1, v1), d20 = dot3(v2, v0), d21 = dot3(v2, v1);
  float den = d00 * d11 - d01 * d01;
  float lv = den != 0.f ? (d11 * d20 - d01 * d21) / den : 0.f, lw = den != 0.f ? (d00 * d21 - d01 * d20) / den : 0.f;
  float lu = 1.f - lv - lw;
  float pa[3], pb[3];
  for (int q = 0; q < 3; q++) {
    pa[q] = lu * V[pv[0]].a[q] + lv * V[pv[1]].a[q] + lw * V[pv[2]].a[q];
    pb[q] = lu * V[pv[0]].b[q] + lv * V[pv[1]].b[q] + lw * V[pv[2]].b[q];
  }
  // the supports along the normal give the depth h_A(n) + h_B(-n) to second
  // order in the normal's error; the points move onto the supporting planes
  KSv sw;
  k_msupport(A, B, pn, &sw);
  float ha = dot3(pn, sw.a), hb = dot3(pn, sw.b);
  float ea = ha - dot3(pn, pa), eb = hb - dot3(pn, pb);
  out->dist = -(ha - hb);
  for (int q = 0; q < 3; q++) {
    out->n[q] = pn[q];
    out->pos[q] = 0.5f * (pa[q] + ea * pn[q] + pb[q] + eb * pn[q]);
  }
  return 1;
}

// narrow phase for one pair (geom1 has the lower type); -1 for unsupported
template <int NPSET>
KDEV int k_narrow(int t1, int t2, const float* s1, const float* s2, const float* x1, const float* R1,
                  const float* x2, const float* R2, float margin, RawC* out) {
  if (t1 == GPLANE) {
    float pn[3] = {R1[2], R1[5], R1[8]};
    if (t2 == GSPHERE) return k_plane_point(x1, pn, x2, s2[0], margin, out);
    if (NPSET >= 2 && (t2 == GELLIPSOID || t2 == GCYLINDER)) {  // the deepest support point
      KCvx g = {t2, x2, R2, s2};
      float dn[3] = {-pn[0], -pn[1], -pn[2]}, sp[3];
      k_support(g, dn, sp);
      return k_plane_point(x1, pn, sp, 0.f, margin, out);
    }
    if (t2 == GCAPSULE) {
      float p0[3], dd[3];
      k_capsule_ends(x2, R2, s2[1], p0, dd);
      float pA[3] = {p0[0] + dd[0], p0[1] + dd[1], p0[2] + dd[2]};
      int n = k_plane_point(x1, pn, pA, s2[0], margin, out);
      n += k_plane_point(x1, pn, p0, s2[0], margin, out + n);
      return n;
    }
    if (t2 == GBOX) {
      int n = 0;
      for (int k = 0; k < 8 && n < 4; k++) {
        float cl[3] = {(k & 1) ? s2[0] : -s2[0], (k & 2) ? s2[1] : -s2[1], (k & 4) ? s2[2] : -s2[2]}, cw[3];
        mv3(R2, cl, cw);
        cw[0] += x2[0]; cw[1] += x2[1]; cw[2] += x2[2];
        n += k_plane_point(x1, pn, cw, 0.f, margin, out + n);
      }
      return n;
    }
    return -1;
  }
  if (NPSET >= 2 && (t1 == GELLIPSOID || t1 == GCYLINDER || t2 == GELLIPSOID || t2 == GCYLINDER)) {
    KCvx A = {t1, x1, R1, s1}, B = {t2, x2, R2, s2};
    return k_gjk_epa(A, B, margin, out);
  }
  if (t1 == GSPHERE) {
    if (t2 == GSPHERE) return k_sphere_sphere(x1, s1[0], x2, s2[0], margin, out);
    if (t2 == GCAPSULE) {
      float p0[3], dd[3];
      k_capsule_ends(x2, R2, s2[1], p0, dd);
      float rel[3] = {x1[0] - p0[0], x1[1] - p0[1], x1[2] - p0[2]};
      float a = dot3(dd, dd);
      float t = a > KMINVAL ? clampf(dot3(rel, dd) / a, 0.f, 1.f) : 0.f;
      float c2[3] = {p0[0] + t * dd[0], p0[1] + t * dd[1], p0[2] + t * dd[2]};
      return k_sphere_sphere(x1, s1[0], c2, s2[0], margin, out);
    }
    if (t2 == GBOX) return k_sphere_box(x1, s1[0], x2, R2, s2, margin, out);
    return -1;
  }
  if (NPSET >= 1 && t1 == GCAPSULE && t2 == GBOX) return k_capsule_box(x1, R1, s1, x2, R2, s2, margin, out);
  if (NPSET >= 1 && t1 == GBOX && t2 == GBOX) return k_box_box(x1, R1, s1, x2, R2, s2, margin, out);
  if (t1 == GCAPSULE && t2 == GCAPSULE) {
    float a0[3], ad[3], b0[3], bd[3], s, t;
    k_capsule_ends(x1, R1, s1[1], a0, ad);
    k_capsule_ends(x2, R2, s2[1], b0, bd);
    k_seg_seg(a0, ad, b0, bd, &s, &t);
    float c1[3] = {a0[0] + s * ad[0], a0[1] + s * ad[1], a0[2] + s * ad[2]};
    float c2[3] = {b0[0] + t * bd[0], b0[1] + t * bd[1], b0[2] + t * bd[2]};
    return k_sphere_sphere(c1, s1[0], c2, s2[0], margin, out);
  }
  return -1;
}

// contact frame from the normal (MuJoCo mju_makeFrame)
KDEV void k_make_frame(const float* n_in, float* f) {
  float n[3] = {n_in[0], n_in[1], n_in[2]};
  float nn = sqrtf(dot3(n, n));
  if (nn < KMINVAL) { n[0] = 1.f; n[1] = n[2] = 0.f; }
  else { float in = 1.0f / nn; n[0] *= in; n[1] *= in; n[2] *= in; }
  float y[3];
  if (fabsf(n[1]) < 0.5f) { y[0] = 0.f; y[1] = 1.f; y[2] = 0.f; }
  else { y[0] = 0.f; y[1] = 0.f; y[2] = 1.f; }
  float s = dot3(n, y);
  for (int i = 0; i < 3; i++) y[i] -= s * n[i];
  float yn = sqrtf(dot3(y, y));
  float iy = 1.0f / yn;
  y[0] *= iy; y[1] *= iy; y[2] *= iy;
  float z[3];
  cross3(n, y, z);
  f[0] = n[0]; f[1] = n[1]; f[2] = n[2];
  f[3] = y[0]; f[4] = y[1]; f[5] = y[2];
  f[6] = z[0]; f[7] = z[1]; f[8] = z[2];
}

// mj_collision over the static pair table: bounding-sphere cull
// (reference collision.hpp:301-305), narrow phase per lane, wave prefix scan
// to compact contacts in pair order.
template <int NPSET>
KDEV int st_collision(const DModel& m, const Layout& L, float* w) {
  int* misc = iptr(w, L.misc);
  int st = 0;
  if (m.disableflags & (DS_CONTACT | DS_CONSTRAINT)) {
    if (LANE == 0) { misc[MISC_NCON] = 0; misc[MISC_NCON_ALL] = 0; }
    KSYNC();
    return 0;
  }
  const float *gx = w + L.gxpos, *gm = w + L.gxmat;
  float* con = w + L.con;
  int base = 0;
  for (int p0 = 0; p0 < m.npair; p0 += WAVE) {
    int p = p0 + LANE;
    RawC rc[MAXC_PAIR];
    int cnt = 0;
    if (p < m.npair) {
      int g1 = m.pair_g1[p], g2 = m.pair_g2[p];
      const float *x1 = gx + 3 * g1, *x2 = gx + 3 * g2;
      float rb1 = m.geom_rbound[g1], rb2 = m.geom_rbound[g2];
      float margin = m.pair_margin[p];
      bool cull = false;
      if (rb1 > 0.f && rb2 > 0.f) {
        float dv[3] = {x1[0] - x2[0], x1[1] - x2[1], x1[2] - x2[2]};
        cull = sqrtf(dot3(dv, dv)) > rb1 + rb2 + margin;
      }
      if (!cull) cnt = k_narrow<NPSET>(m.geom_type[g1], m.geom_type[g2], m.geom_size + 3 * g1, m.geom_size + 3 * g2,
                                x1, gm + 9 * g1, x2, gm + 9 * g2, margin, rc);
    }
    int badp = cnt < 0 ? 1 : 0;
    if (cnt < 0) cnt = 0;
    if (wori(badp)) st |= ST_BAD_PAIR;
    int off = wscan_excl(cnt);
    int tot = wsumi(cnt);
    for (int k = 0; k < cnt; k++) {
      int idx = base + off + k;
      if (idx >= L.maxcon) break;
      float* c = con + CON_STRIDE * idx;
      c[CON_POS] = rc[k].pos[0]; c[CON_POS + 1] = rc[k].pos[1]; c[CON_POS + 2] = rc[k].pos[2];
      k_make_frame(rc[k].n, c + CON_FRAME);
      c[CON_DIST] = rc[k].dist;
      iptr(c, CON_PAIR)[0] = p;
    }
    base += tot;
  }
  // all contacts count up to con_cap (the reference's list cap); the first
  // maxcon, every one the row budget can admit, are stored
  int all = base;
  if (all > L.con_cap) { st |= ST_CON_OVERFLOW; all = L.con_cap; }
  if (all > L.maxcon) st |= ST_EFC_OVERFLOW;  // unstored contacts are past the row budget
  if (LANE == 0) { misc[MISC_NCON] = base < L.maxcon ? base : L.maxcon; misc[MISC_NCON_ALL] = all; }
  KSYNC();
  return st;
}

// ------------------------------------------------------------------ constraints
KDEV float k_impedance(const float* si, float x) {
  float dmin = clampf(si[0], KMINIMP, KMAXIMP), dmax = clampf(si[1], KMINIMP, KMAXIMP);
  float width = si[2], mid = clampf(si[3], KMINIMP, KMAXIMP), power = si[4] < 1.f ? 1.f : si[4];
  if (dmin == dmax || width <= KMINVAL) return dmax;
  float xx = fabsf(x) / width;
  if (xx >= 1.f) return dmax;
  float y;
  if (power == 1.f) y = xx;
  else if (xx <= mid) y = powf(xx, power) / powf(mid, power - 1.f);
  else y = 1.f - powf(1.f - xx, power) / powf(1.f - mid, power - 1.f);
  return dmin + y * (dmax - dmin);
}

// mj_makeConstraint + impedance + reference acceleration: dof friction rows,
// joint-limit rows (one per violated side), contact rows (frictionless or
// pyramidal). Row layout and formulas match the FP64 oracle; the
// per-row quantities follow reference solver.hpp:120-139,299-371,398-587.
KDEV void build_body_contacts(const DModel& m, const Layout& L, float* w);

KDEV int st_constraints(const DModel& m, const Layout& L, float* w) {
  int* misc = iptr(w, L.misc);
  int nv = m.nv;
  int st = 0;
  float *J = w + L.efc_J, *D = w + L.efc_D, *aref = w + L.efc_aref;
  float *epos = w + L.efc_pos, *emarg = w + L.efc_margin;
  int *etype = iptr(w, L.efc_type), *eid = iptr(w, L.efc_id);
  const float *qpos = w + L.qpos, *qvel = w + L.qvel;
  float* con = w + L.con;
  int nrow = 0;
  if (m.disableflags & DS_CONSTRAINT) {
    if (LANE == 0) misc[MISC_NEFC] = 0;
    KSYNC();
    return 0;
  }
  // dof friction rows
  // row budget (reference src/compat/mkx_kernels/solver.hpp:299,349,403):
  // friction and limit rows while fewer than efc_cap - 4 rows exist, a
  // contact while fewer than efc_cap - 10
  int lim_cap = L.efc_cap - 4 < L.maxefc ? L.efc_cap - 4 : L.maxefc;
  if (lim_cap < 0) lim_cap = 0;
  int nf = (m.disableflags & DS_FRICTIONLOSS) ? 0 : m.nfloss;
  if (nf > lim_cap) { nf = lim_cap; st |= ST_EFC_OVERFLOW; }
  for (int t = LANE; t < nf; t += WAVE) {
    int k = m.floss_dof[t];
    etype[t] = EFC_FRICTION_DOF; eid[t] = k; epos[t] = 0.f; emarg[t] = 0.f;
  }
  nrow = nf;
  // joint limits: 0..2 rows per limited joint
  if (!(m.disableflags & DS_LIMIT)) {
    for (int t0 = 0; t0 < m.nlim; t0 += WAVE) {
      int t = t0 + LANE;
      int cnt = 0, j = -1;
      float d0 = 0.f, d1 = 0.f;
      bool s0 = false, s1 = false;
      if (t < m.nlim) {
        j = m.lim_jnt[t];
        float v = qpos[m.jnt_qposadr[j]];
        d0 = v - m.jnt_range[2 * j];
        d1 = m.jnt_range[2 * j + 1] - v;
        s0 = d0 < m.jnt_margin[j];
        s1 = d1 < m.jnt_margin[j];
        cnt = (int)s0 + (int)s1;
      }
      int off = wscan_excl(cnt), tot = wsumi(cnt);
      int r = nrow + off;
      if (s0 && r < lim_cap) { etype[r] = EFC_LIMIT_JOINT; eid[r] = 2 * j; epos[r] = d0; emarg[r] = m.jnt_margin[j]; }
      if (s0) r++;
      if (s1 && r < lim_cap) { etype[r] = EFC_LIMIT_JOINT; eid[r] = 2 * j + 1; epos[r] = d1; emarg[r] = m.jnt_margin[j]; }
      nrow += tot;
    }
    if (nrow > lim_cap) { nrow = lim_cap; st |= ST_EFC_OVERFLOW; }
  }
  // contact rows: count per contact, compact
  int ncon = misc[MISC_NCON];
  if (!(m.disableflags & DS_CONTACT)) {
    int cut = 0x7fffffff;
    for (int c0 = 0; c0 < ncon; c0 += WAVE) {
      int c = c0 + LANE;
      int cnt = 0;
      if (c < ncon) {
        float* cr = con + CON_STRIDE * c;
        int p = iptr(cr, CON_PAIR)[0];
        float incl = m.pair_margin[p] - m.pair_gap[p];
        int dim = m.pair_condim[p];
        if (cr[CON_DIST] < incl) cnt = dim == 1 ? 1 : (m.cone == 0 ? 2 * (dim - 1) : dim);
      }
      int off = wscan_excl(cnt), tot = wsumi(cnt);
      int dropr0 = 0x7fffffff;
      if (c < ncon) {
        float* cr = con + CON_STRIDE * c;
        int r0 = nrow + off;
        if (cnt > 0 && (r0 >= L.efc_cap - 10 || r0 + cnt > L.maxefc)) { dropr0 = r0; cnt = 0; }
        iptr(cr, CON_NROW)[0] = cnt;
        iptr(cr, CON_ROW0)[0] = r0;
        int p = iptr(cr, CON_PAIR)[0];
        int dim = m.pair_condim[p];
        for (int k = 0; k < cnt; k++) {
          etype[r0 + k] = dim == 1 ? EFC_CONTACT_FRICTIONLESS : (m.cone == 0 ? EFC_CONTACT_PYRAMIDAL : EFC_CONTACT_ELLIPTIC);
          bool fric = m.cone != 0 && dim > 1 && k > 0;  // elliptic friction rows carry no position term
          eid[r0 + k] = c; epos[r0 + k] = fric ? 0.f : cr[CON_DIST];
          emarg[r0 + k] = fric ? 0.f : m.pair_margin[p] - m.pair_gap[p];
        }
      }
      nrow += tot;
      int dmin = wmini(dropr0);
      if (dmin < cut) cut = dmin;
    }
    if (cut < nrow) { nrow = cut; st |= ST_EFC_OVERFLOW; }
    if (nrow > L.maxefc) { nrow = L.maxefc; st |= ST_EFC_OVERFLOW; }
  }
  KSYNC();
  // Jacobian rows: dense (nv entries per row) or, beyond SPARSE_ROWS_NV dofs,
  // compressed (jw entries per row with their dof indices)
  const int jw = MJK_BIG ? L.jw : 0, rw = jw > 0 ? jw : nv;
  int* Jidx = jw > 0 ? iptr(w, L.efc_Jidx) : nullptr;
  for (int t = LANE; t < nrow * rw; t += WAVE) {
    J[t] = 0.f;
    if (jw > 0) Jidx[t] = -1;
  }
  KSYNC();
  int nlr = nrow;
  for (int r = LANE; r < nlr; r += WAVE) {
    int ty = etype[r];
    int k = -1;
    float v = 0.f;
    if (ty == EFC_FRICTION_DOF) { k = eid[r]; v = 1.f; }
    else if (ty == EFC_LIMIT_JOINT) {
      int j = eid[r] >> 1;
      k = m.jnt_dofadr[j];
      v = (eid[r] & 1) ? -1.f : 1.f;
    }
    if (k >= 0) {
      if (jw > 0) { J[r * jw] = v; Jidx[r * jw] = k; }
      else J[r * nv + k] = v;
    }
  }
  const float *cdof = w + L.cdof, *scom = w + L.scom;
  // contacts x entries: entry k is dof k (dense) or the contact's k-th chain
  // dof (compressed: body2's chain from its last dof up, then body1's up to
  // the first dof on body2's chain)
  for (int t = LANE; t < ncon * rw; t += WAVE) {
    int c = t / rw, k = t - c * rw;
    float* cr = con + CON_STRIDE * c;
    int cnt = iptr(cr, CON_NROW)[0];
    if (cnt == 0) continue;
    int r0 = iptr(cr, CON_ROW0)[0];
    int p = iptr(cr, CON_PAIR)[0];
    int b1 = m.pair_b1[p], b2 = m.pair_b2[p];
    int slot = k;
    if (jw > 0) {
      int l2 = m.body_lastdof[b2], l1 = m.body_lastdof[b1];
      int n2 = l2 >= 0 ? m.dof_nchain[l2] : 0;
      if (slot < n2) k = m.dof_chain[l2 * MAX_DOF_DEPTH + slot];
      else {
        // the q-th dof of body1's chain off body2's chain (a prefix of it)
        int q = slot - n2, d = l1;
        while (d >= 0) {
          if (l2 >= d && l2 < m.dof_subtree_end[d]) { d = -1; break; }  // the shared part
          if (q == 0) break;
          q--;
          d = m.dof_parentid[d];
        }
        if (d < 0) continue;  // past the row's entries
        k = d;
      }
    }
    float jp[3] = {0.f, 0.f, 0.f}, jr[3] = {0.f, 0.f, 0.f};
    const float* cd = cdof + 6 * k;
    // dof k moves body b iff it is in the chain of b's last dof: k <= d < subtree_end(k)
    int d2 = m.body_lastdof[b2], d1 = m.body_lastdof[b1];
    if (d2 >= k && d2 < m.dof_subtree_end[k]) {
      const float* cc = scom + 3 * m.body_rootid[b2];
      float off[3] = {cr[0] - cc[0], cr[1] - cc[1], cr[2] - cc[2]}, tt[3];
      cross3(cd, off, tt);
      for (int i = 0; i < 3; i++) { jp[i] += cd[3 + i] + tt[i]; jr[i] += cd[i]; }
    }
    if (d1 >= k && d1 < m.dof_subtree_end[k]) {
      const float* cc = scom + 3 * m.body_rootid[b1];
      float off[3] = {cr[0] - cc[0], cr[1] - cc[1], cr[2] - cc[2]}, tt[3];
      cross3(cd, off, tt);
      for (int i = 0; i < 3; i++) { jp[i] -= cd[3 + i] + tt[i]; jr[i] -= cd[i]; }
    }
    const float* fr = cr + CON_FRAME;
    float Jn = dot3(fr, jp);
    // entry (row r, slot) of the dense or compressed layout
    auto at = [&](int r) -> float& { return J[r * rw + slot]; };
    if (jw > 0)
      for (int q = 0; q < cnt; q++) Jidx[(r0 + q) * jw + slot] = k;
    if (cnt == 1) { at(r0) = Jn; continue; }
    float Jd[5] = {dot3(fr + 3, jp), dot3(fr + 6, jp), dot3(fr, jr), dot3(fr + 3, jr), dot3(fr + 6, jr)};
    if (m.cone != 0) {  // elliptic: normal, then t1, t2, spin, roll1, roll2
      at(r0) = Jn;
      for (int q = 1; q < cnt; q++) at(r0 + q) = Jd[q - 1];
      continue;
    }
    const float* mu = m.pair_friction + 5 * p;
    float sc = fric_scale(m, L, w);
    for (int q = 0; q < cnt / 2; q++) {
      float mq = mu[q] * sc;
      at(r0 + 2 * q) = Jn + mq * Jd[q];
      at(r0 + 2 * q + 1) = Jn - mq * Jd[q];
    }
  }
  KSYNC();
  // impedance, aref, regulariser
  for (int r = LANE; r < nrow; r += WAVE) {
    int ty = etype[r];
    const float *sr, *si;
    float dA;
    float mu0 = 0.f;
    if (ty == EFC_FRICTION_DOF) {
      int k = eid[r];
      sr = m.dof_solref + 2 * k; si = m.dof_solimp + 5 * k; dA = m.dof_invweight0[k];
    } else if (ty == EFC_LIMIT_JOINT) {
      int j = eid[r] >> 1;
      sr = m.jnt_solref + 2 * j; si = m.jnt_solimp + 5 * j; dA = m.dof_invweight0[m.jnt_dofadr[j]];
    } else {
      int p = iptr(con + CON_STRIDE * eid[r], CON_PAIR)[0];
      sr = m.pair_solref + 2 * p; si = m.pair_solimp + 5 * p;
      float tran = m.body_invweight0[2 * m.pair_b1[p]] + m.body_invweight0[2 * m.pair_b2[p]];
      mu0 = m.pair_friction[5 * p] * fric_scale(m, L, w);
      dA = ty == EFC_CONTACT_PYRAMIDAL ? tran * (1.f + mu0 * mu0) : tran;
    }
    float v = 0.f;
    if (jw > 0) {
      for (int q = 0; q < jw; q++) {
        int k = Jidx[r * jw + q];
        if (k < 0) break;
        v += J[r * jw + q] * qvel[k];
      }
    } else {
      v = kdot(J + r * nv, 1, qvel, 1, nv);
    }
    float tc = sr[0], dr = sr[1];
    float dmax = clampf(si[1], KMINIMP, KMAXIMP);
    float kk, bb;
    if (tc > 0.f) {
      if (!(m.disableflags & DS_REFSAFE)) tc = fmaxf(tc, 2.f * m.timestep);
      kk = 1.f / (dmax * dmax * tc * tc * dr * dr);
      bb = 2.f / (dmax * tc);
    } else {
      kk = -tc / (dmax * dmax);
      bb = -dr / dmax;
    }
    float x = epos[r] - emarg[r];
    float imp = k_impedance(si, x);
    aref[r] = -bb * v - kk * imp * x;
    float R = fmaxf(dA * (1.f - imp) / imp, KMINVAL);
    if (ty == EFC_CONTACT_PYRAMIDAL) R = fmaxf(2.f * mu0 * mu0 / m.impratio * R, KMINVAL);
    if (ty == EFC_CONTACT_ELLIPTIC) {
      // friction rows (MuJoCo mj_makeImpedance): R_1 = R_0 / impratio,
      // R_k = R_1 mu_0^2 / mu_(k-1)^2, R_0 the normal row's regulariser
      const float* cr = con + CON_STRIDE * eid[r];
      int kr = r - iptr(cr, CON_ROW0)[0];
      if (kr > 0) {
        int p = iptr(cr, CON_PAIR)[0];
        float imp0 = k_impedance(si, cr[CON_DIST] - (m.pair_margin[p] - m.pair_gap[p]));
        float R0 = fmaxf(dA * (1.f - imp0) / imp0, KMINVAL);
        float R1 = R0 / fmaxf(KMINVAL, m.impratio);
        const float* f5 = m.pair_friction + 5 * p;
        float sc = fric_scale(m, L, w), f0 = f5[0] * sc, fk = f5[kr - 1] * sc;
        R = fmaxf(kr == 1 ? R1 : R1 * f0 * f0 / (fk * fk), KMINVAL);
      }
    }
    D[r] = 1.f / R;  // friction rows take R back as 1 / D
  }
  if (LANE == 0) misc[MISC_NEFC] = nrow;
  KSYNC();
  if (jw > 0) build_body_contacts(m, L, w);
  return st;
}

// ------------------------------------------------------------------ solver
// y = M x (dense qM), lanes over rows
// row i of M x. Up to 64 dofs the dense row (LDS-resident, no dependent
// loads); beyond, only the dof tree's nonzeros: ancestors (the parent chain),
// the diagonal and the descendants (a contiguous dof range)
KDEV float mulM_row(const DModel& m, const float* M, const float* x, int i) {
  int nv = m.nv;
  if (!m_compact(m)) return kdot(M + i * nv, 1, x, 1, nv);
  float acc = 0.f;
  const float* row = M + m.dof_madr[i];  // the diagonal, then the ancestors
  int a = 1;
  for (int j = m.dof_parentid[i]; j >= 0; j = m.dof_parentid[j], a++) acc += row[a] * x[j];
  for (int k = i; k < m.dof_subtree_end[i]; k++) acc += M[mxi(m, k, i)] * x[k];
  return acc;
}

KDEV void mulM(const DModel& m, const float* M, const float* x, float* y) {
  if (path_trees(m)) {
    mulM_trees(m, M, x, y);
    return;
  }
  for (int i = LANE; i < m.nv; i += WAVE) y[i] = mulM_row(m, M, x, i);
}

// row r of J x. Beyond 64 dofs only the row's possible nonzeros: a friction
// or limit row's dof, a contact row's two body chains (body2's, then body1's
// up to the first dof body2's chain shares); rows of other types are dense.
KDEV float jrow_dot(const DModel& m, const Layout& L, const float* w, int r, const float* x) {
  int nv = m.nv;
  if (MJK_BIG && L.jw > 0) {  // compressed row: its entries and their dofs
    const float* Jr = w + L.efc_J + r * L.jw;
    const int* ix = iptr(w, L.efc_Jidx) + r * L.jw;
    float acc = 0.f;
    for (int q = 0; q < L.jw; q++) {
      int k = ix[q];
      if (k < 0) break;
      acc += Jr[q] * x[k];
    }
    return acc;
  }
  const float* Jr = w + L.efc_J + r * nv;
  if (!MJK_BIG || nv <= 64) return kdot(Jr, 1, x, 1, nv);
  int ty = reinterpret_cast<const int*>(w + L.efc_type)[r], id = reinterpret_cast<const int*>(w + L.efc_id)[r];
  if (ty == EFC_FRICTION_DOF) return Jr[id] * x[id];
  if (ty == EFC_LIMIT_JOINT) { int d = m.jnt_dofadr[id >> 1]; return Jr[d] * x[d]; }
  if (ty != EFC_CONTACT_FRICTIONLESS && ty != EFC_CONTACT_PYRAMIDAL && ty != EFC_CONTACT_ELLIPTIC)
    return kdot(Jr, 1, x, 1, nv);
  int p = reinterpret_cast<const int*>(w + L.con + CON_STRIDE * id + CON_PAIR)[0];
  int l2 = m.body_lastdof[m.pair_b2[p]];
  float acc = 0.f;
  for (int d = l2; d >= 0; d = m.dof_parentid[d]) acc += Jr[d] * x[d];
  for (int d = m.body_lastdof[m.pair_b1[p]]; d >= 0; d = m.dof_parentid[d]) {
    if (l2 >= d && l2 < m.dof_subtree_end[d]) break;  // d is on body2's chain
    acc += Jr[d] * x[d];
  }
  return acc;
}

// y[r] = J_r x - sub[r] (sub may be null) for every row. Compressed rows
// (in device memory beyond 144 KB workspaces) go two per lane at a time with
// every entry's dof index and value loaded before the gathers of x: the
// loads of a row do not wait for each other (a row's entries past its end,
// index -1, add 0 exactly).
KDEV void jrows_dot(const DModel& m, const Layout& L, const float* w, int nefc, const float* x, float* y,
                    const float* sub) {
  if (MJK_BIG && L.jw > 0) {
    const int jw = L.jw;
    const float* J = w + L.efc_J;
    const int* ix = iptr(w, L.efc_Jidx);
    for (int r = LANE; r < nefc; r += 2 * WAVE) {
      int r2 = r + WAVE < nefc ? r + WAVE : r;
      const float *Ja = J + r * jw, *Jb = J + r2 * jw;
      const int *ia = ix + r * jw, *ib = ix + r2 * jw;
      float a = 0.f, b = 0.f;
#pragma unroll 4
      for (int q = 0; q < jw; q++) {
        int ka = ia[q], kb = ib[q];
        a = fmaf(ka >= 0 ? Ja[q] : 0.f, x[ka >= 0 ? ka : 0], a);
        b = fmaf(kb >= 0 ? Jb[q] : 0.f, x[kb >= 0 ? kb : 0], b);
      }
      y[r] = sub ? a - sub[r] : a;
      if (r + WAVE < nefc) y[r + WAVE] = sub ? b - sub[r + WAVE] : b;
    }
    return;
  }
  for (int r = LANE; r < nefc; r += WAVE) {
    float a = jrow_dot(m, L, w, r, x);
    y[r] = sub ? a - sub[r] : a;
  }
}

// Per-body contact lists for mul_jt_sparse, built once per substep after the
// rows: body b's entries are 2 c + side (side 0: b is the contact's body2,
// 1: its body1) in contact order, CSR in bcon_adr / bcon.
KDEV void build_body_contacts(const DModel& m, const Layout& L, float* w) {
  int nb = m.nbody;
  int ncon = iptr(w, L.misc)[MISC_NCON];
  const float* con = w + L.con;
  int *adr = iptr(w, L.bcon_adr), *lst = iptr(w, L.bcon);
  int* cw = iptr(w, L.cwrench);  // body pairs of the contacts, staged in the wrench scratch
  for (int c = LANE; c < ncon; c += WAVE) {
    const float* cr = con + CON_STRIDE * c;
    int p = iptr(cr, CON_PAIR)[0];
    bool rows = iptr(cr, CON_NROW)[0] > 0;
    cw[2 * c] = rows ? m.pair_b2[p] : -1;
    cw[2 * c + 1] = rows ? m.pair_b1[p] : -1;
  }
  KSYNC();
  int run = 0;
  for (int b0 = 0; b0 < nb; b0 += WAVE) {
    int b = b0 + LANE, cnt = 0;
    if (b < nb && b > 0)
      for (int t = 0; t < 2 * ncon; t++) cnt += cw[t] == b;
    int off = run + wscan_excl(cnt);
    if (b < nb) adr[b] = off;
    run += wsumi(cnt);
    if (b < nb && b > 0) {
      int k = off;
      for (int t = 0; t < 2 * ncon; t++)
        if (cw[t] == b) lst[k++] = t;
    }
  }
  if (LANE == 0) adr[nb] = run;
  KSYNC();
}

// y = J^T f for compressed rows, without the rows: each contact's rows add
// up to a force F and torque T at the contact point (normal n, tangents t1, t2
// of its frame; pyramidal edge pairs J_n +- mu_q J_q, elliptic rows n, t1, t2,
// then spin and rolls as torques), applied +/- to body2 / body1 as a wrench
// about the tree root's subtree COM (the point cdof refers to); a dof then
// takes cdof . (sum of the wrenches on its body's subtree). Friction and
// limit rows (the leading rows) add J f to their dof. One lane per contact
// for its two wrenches, one lane per body over its contact list (in contact
// order), one lane per dof: deterministic sums, work ncon + list lengths +
// nv x subtree size instead of nefc x nv.
KDEV void mul_jt_sparse(const DModel& m, const Layout& L, const float* w, const float* f, int nefc, float* y) {
  int nv = m.nv, nb = m.nbody;
  float* wr = const_cast<float*>(w) + L.wrench;
  float* cw = const_cast<float*>(w) + L.cwrench;
  const int* misc = iptr(w, L.misc);
  int ncon = misc[MISC_NCON];
  const float *con = w + L.con, *scom = w + L.scom, *cdof = w + L.cdof;
  float sc = fric_scale(m, L, w);
  for (int c = LANE; c < ncon; c += WAVE) {
    const float* cr = con + CON_STRIDE * c;
    int cnt = iptr(cr, CON_NROW)[0];
    if (cnt == 0) continue;
    int p = iptr(cr, CON_PAIR)[0];
    const float* fr = f + iptr(cr, CON_ROW0)[0];
    const float *n = cr + CON_FRAME, *t1 = n + 3, *t2 = n + 6;
    float F[3], T[3] = {0.f, 0.f, 0.f};
    if (cnt == 1) {
      for (int i = 0; i < 3; i++) F[i] = fr[0] * n[i];
    } else if (m.cone != 0) {
      for (int i = 0; i < 3; i++) F[i] = fr[0] * n[i] + fr[1] * t1[i] + fr[2] * t2[i];
      if (cnt > 3) for (int i = 0; i < 3; i++) T[i] = fr[3] * n[i];
      if (cnt > 4) for (int i = 0; i < 3; i++) T[i] += fr[4] * t1[i] + fr[5] * t2[i];
    } else {
      const float* mu = m.pair_friction + 5 * p;
      for (int i = 0; i < 3; i++) F[i] = 0.f;
      for (int q = 0; q < cnt / 2; q++) {
        float fp = fr[2 * q], fm = fr[2 * q + 1], dd = mu[q] * sc * (fp - fm);
        const float* dir = q == 0 ? t1 : (q == 1 ? t2 : (q == 2 ? n : (q == 3 ? t1 : t2)));
        float* acc = q < 2 ? F : T;
        for (int i = 0; i < 3; i++) { F[i] += (fp + fm) * n[i]; acc[i] += dd * dir[i]; }
      }
    }
    for (int side = 0; side < 2; side++) {
      int b = side == 0 ? m.pair_b2[p] : m.pair_b1[p];
      float sg = side == 0 ? 1.f : -1.f;
      const float* cc = scom + 3 * m.body_rootid[b];
      float off[3] = {cr[0] - cc[0], cr[1] - cc[1], cr[2] - cc[2]}, t[3];
      cross3(off, F, t);
      float* o = cw + 12 * c + 6 * side;
      for (int i = 0; i < 3; i++) { o[i] = sg * (t[i] + T[i]); o[3 + i] = sg * F[i]; }
    }
  }
  KSYNC();
  const int *adr = iptr(w, L.bcon_adr), *lst = iptr(w, L.bcon);
  for (int b = LANE; b < nb; b += WAVE) {
    float W[6] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
    for (int e = adr[b]; e < adr[b + 1]; e++) {
      const float* o = cw + 6 * lst[e];  // 12 c + 6 side
      for (int i = 0; i < 6; i++) W[i] += o[i];
    }
    for (int i = 0; i < 6; i++) wr[6 * b + i] = W[i];
  }
  KSYNC();
  const float* J = w + L.efc_J;
  const int *ix = iptr(w, L.efc_Jidx), *etype = iptr(w, L.efc_type);
  for (int d = LANE; d < nv; d += WAVE) {
    const float* cd = cdof + 6 * d;
    int bd = m.dof_bodyid[d];
    float acc = 0.f;
    for (int bb = bd; bb < m.body_subtree_end[bd]; bb++) {
      const float* W = wr + 6 * bb;
      acc += cd[0] * W[0] + cd[1] * W[1] + cd[2] * W[2] + cd[3] * W[3] + cd[4] * W[4] + cd[5] * W[5];
    }
    for (int r = 0; r < nefc; r++) {
      int ty = etype[r];
      if (ty != EFC_FRICTION_DOF && ty != EFC_LIMIT_JOINT) break;
      if (ix[r * L.jw] == d) acc += J[r * L.jw] * f[r];
    }
    y[d] = acc;
  }
}

// y = J^T f. GPU beyond 64 dofs: lane l keeps dofs base + l, base + l + 64,
// .. (four per lane, 256-dof chunks) in registers and streams the rows,
// coalesced across lanes.
KDEV void mul_jt(const DModel& m, const Layout& L, const float* w, const float* f, int nefc, float* y) {
  int nv = m.nv;
  if (MJK_BIG && L.jw > 0) {
    mul_jt_sparse(m, L, w, f, nefc, y);
    return;
  }
  const float* J = w + L.efc_J;
#if WAVE == 64
  if (MJK_BIG && nv > WAVE) {
    for (int base = 0; base < nv; base += 4 * WAVE) {
      // eight rows per step: up to 32 independent loads in flight per lane
      float g[4] = {0.f, 0.f, 0.f, 0.f};
      int r = 0;
      for (; r + 8 <= nefc; r += 8) {
        float fr[8];
#pragma unroll
        for (int u = 0; u < 8; u++) fr[u] = f[r + u];
#pragma unroll
        for (int q = 0; q < 4; q++) {
          int i = base + LANE + q * WAVE;
          if (i < nv) {
            float a[8];
#pragma unroll
            for (int u = 0; u < 8; u++) a[u] = J[(r + u) * nv + i];
#pragma unroll
            for (int u = 0; u < 8; u++) g[q] += a[u] * fr[u];
          }
        }
      }
      for (; r < nefc; r++) {
        float fr = f[r];
#pragma unroll
        for (int q = 0; q < 4; q++)
          if (base + LANE + q * WAVE < nv) g[q] += J[r * nv + base + LANE + q * WAVE] * fr;
      }
#pragma unroll
      for (int q = 0; q < 4; q++)
        if (base + LANE + q * WAVE < nv) y[base + LANE + q * WAVE] = g[q];
    }
    return;
  }
#endif
  for (int i = LANE; i < nv; i += WAVE) y[i] = kdot(J + i, nv, f, 1, nefc);
}

// qacc <- x; Ma = M x; Jaref = J x - aref
// Mx (when not null): M x already known (qfrc_smooth for x = qacc_smooth)
KDEV void set_qacc(const DModel& m, const Layout& L, float* w, const float* x, int nefc, const float* Mx = nullptr) {
  int nv = m.nv;
  float *qacc = w + L.qacc, *Ma = w + L.Ma, *Jaref = w + L.efc_Jaref;
  const float *aref = w + L.efc_aref, *M = w + L.qM;
  if (!Mx && path_trees(m)) mulM_trees(m, M, x, Ma);
  else
    for (int i = LANE; i < nv; i += WAVE) Ma[i] = Mx ? Mx[i] : mulM_row(m, M, x, i);
  jrows_dot(m, L, w, nefc, x, Jaref, aref);
  KSYNC();
  for (int i = LANE; i < nv; i += WAVE) qacc[i] = x[i];
  KSYNC();
}

// constraint forces / active flags / cost at the current Jaref, gradient;
// returns the total cost (wave-uniform)
// Elliptic cone of one contact at row values x (rows r0 .. r0+dim-1; f the
// pair's five friction coefficients, mu = f[0]): with U_0 = mu x_0,
// U_j = f_(j-1) x_j, N = U_0, T = |U_1..|: zone 0 (top, N >= mu T) no force;
// zone 1 (bottom, mu N + T <= 0) every row quadratic; zone 2 (middle) cost
// 1/2 Dm (N - mu T)^2 with Dm = D_0 / (mu^2 (1 + mu^2)) (MuJoCo's primal
// elliptic cone). Writes forces when force != null; returns the zone.
KDEV int k_ellip(const float* x, const float* D, const float* f, float sc, int dim, float* cost, float* force) {
  float mu = f[0] * sc, T = 0.f;
  for (int j = 1; j < dim; j++) { float U = (f[j - 1] * sc) * x[j]; T += U * U; }
  T = sqrtf(T);
  float N = mu * x[0];
  if (N >= mu * T || (T <= 0.f && N >= 0.f)) {
    *cost = 0.f;
    if (force) for (int j = 0; j < dim; j++) force[j] = 0.f;
    return 0;
  }
  if (mu * N + T <= 0.f || (T <= 0.f && N < 0.f)) {
    float c = 0.f;
    for (int j = 0; j < dim; j++) {
      c += 0.5f * D[j] * x[j] * x[j];
      if (force) force[j] = -D[j] * x[j];
    }
    *cost = c;
    return 1;
  }
  float Dm = D[0] / (mu * mu * (1.f + mu * mu)), g = N - mu * T;
  *cost = 0.5f * Dm * g * g;
  if (force) {
    float f0 = -Dm * g * mu;
    force[0] = f0;
    for (int j = 1; j < dim; j++) {
      float fj = f[j - 1] * sc;
      force[j] = -f0 * fj * fj * x[j] / T;
    }
  }
  return 2;
}

// elliptic contact cost along x + alpha v: value, slope, curvature; returns the zone
KDEV int k_ellip_line(const float* x0, const float* v, const float* D, const float* f, float sc, int dim, float alpha,
                      float* c, float* d1, float* d2) {
  float x[6];
  for (int j = 0; j < 6; j++) x[j] = j < dim ? x0[j] + alpha * v[j] : 0.f;
  int z = k_ellip(x, D, f, sc, dim, c, nullptr);
  *d1 = 0.f; *d2 = 0.f;
  if (z == 1) {
    for (int j = 0; j < dim; j++) { *d1 += D[j] * x[j] * v[j]; *d2 += D[j] * v[j] * v[j]; }
  } else if (z == 2) {
    float mu = f[0] * sc, T = 0.f, TV = 0.f, VV = 0.f;
    for (int j = 1; j < dim; j++) {
      float fj = f[j - 1] * sc;
      float U = fj * x[j], V = fj * v[j];
      T += U * U; TV += U * V; VV += V * V;
    }
    T = sqrtf(T);
    float Tp = TV / T, Tpp = (VV - Tp * Tp) / T;
    float Dm = D[0] / (mu * mu * (1.f + mu * mu));
    float g = mu * x[0] - mu * T, gp = mu * v[0] - mu * Tp, gpp = -mu * Tpp;
    *d1 = Dm * g * gp;
    *d2 = Dm * (gp * gp + g * gpp);
  }
  return z;
}

// grad = Ma - qfrc_smooth - J^T force at the current forces
KDEV void grad_only(const DModel& m, const Layout& L, float* w, int nefc) {
  int nv = m.nv;
  const float *Ma = w + L.Ma, *qfs = w + L.qfrc_smooth;
  float* grad = w + L.grad;
  mul_jt(m, L, w, w + L.efc_force, nefc, grad);
  KSYNC();
  for (int i = LANE; i < nv; i += WAVE) grad[i] = Ma[i] - qfs[i] - grad[i];
  KSYNC();
}

// with_grad false: forces, active flags and cost only (comparing warm-start
// candidates), no gradient
// chmask (when not null): the rows whose activity changed, bit r % 64 of word
// r / 64, for pyramidal / frictionless / limit / friction-loss rows of at most
// 128 rows on 64 lanes (or the serial interpreter); otherwise both words ~0
KDEV float cost_grad(const DModel& m, const Layout& L, float* w, int nefc, int* nchanged, bool with_grad = true,
                     unsigned long long* chmask = nullptr) {
  int nv = m.nv;
  const float *Jaref = w + L.efc_Jaref, *D = w + L.efc_D;
  float *force = w + L.efc_force, *Dact = w + L.efc_Dact;
  int *act = iptr(w, L.efc_act), *etype = iptr(w, L.efc_type);
  const int* eid = iptr(w, L.efc_id);
  float part = 0.f;
  int nch = 0;
  unsigned long long cm[2] = {0ull, 0ull};
  int q = 0;
  for (int r = LANE; r < nefc; r += WAVE, q++) {
    float x = Jaref[r], d = D[r];
    int ty = etype[r];
    if (ty == EFC_CONTACT_ELLIPTIC) continue;  // per contact, below
    int old = act[r];
    if (ty == EFC_EQUALITY) { act[r] = 1; force[r] = -d * x; part += 0.5f * d * x * x; }
    else if (ty == EFC_FRICTION_DOF) {
      float f = m.dof_frictionloss[eid[r]], R = 1.f / d;
      if (x <= -R * f) { act[r] = 0; force[r] = f; part += -f * x - 0.5f * R * f * f; }
      else if (x >= R * f) { act[r] = 0; force[r] = -f; part += f * x - 0.5f * R * f * f; }
      else { act[r] = 1; force[r] = -d * x; part += 0.5f * d * x * x; }
    } else {
      if (x < 0.f) { act[r] = 1; force[r] = -d * x; part += 0.5f * d * x * x; }
      else { act[r] = 0; force[r] = 0.f; }
    }
    bool chg = act[r] != old;
    nch += chg;
    Dact[r] = act[r] ? d : 0.f;
#if WAVE == 64
    unsigned long long b = __ballot(chg);
    if (q < 2) cm[q] = b;
#elif WAVE == 1
    if (chg && r < 128) cm[r >> 6] |= 1ull << (r & 63);
#else
    (void)q;
#endif
  }
  if (m.cone != 0) {
    const int* misc = reinterpret_cast<const int*>(w + L.misc);
    int ncon = misc[MISC_NCON];
    const float* con = w + L.con;
    for (int c = LANE; c < ncon; c += WAVE) {
      const float* cr = con + CON_STRIDE * c;
      int dim = reinterpret_cast<const int*>(cr + CON_NROW)[0], r0 = reinterpret_cast<const int*>(cr + CON_ROW0)[0];
      if (dim < 2 || r0 + dim > nefc) continue;
      const float* f5 = m.pair_friction + 5 * reinterpret_cast<const int*>(cr + CON_PAIR)[0];
      float cst;
      int z = k_ellip(Jaref + r0, D + r0, f5, fric_scale(m, L, w), dim, &cst, force + r0);
      part += cst;
      // the middle zone's Hessian moves with qacc: it always asks for a new factor
      nch += (act[r0] != z) || z == 2;
      for (int j = 0; j < dim; j++) {
        act[r0 + j] = z;
        Dact[r0 + j] = z == 1 ? D[r0 + j] : 0.f;
      }
    }
  }
  const float *Ma = w + L.Ma, *qfs = w + L.qfrc_smooth, *qacc = w + L.qacc, *qas = w + L.qacc_smooth;
  for (int i = LANE; i < nv; i += WAVE) part += 0.5f * (Ma[i] - qfs[i]) * (qacc[i] - qas[i]);
  float cost = wsum(part);
  *nchanged = wsumi(nch);
  if (chmask) {
    bool known = (WAVE == 64 || WAVE == 1) && nefc <= 128 && m.cone == 0;
#if WAVE == 64
    // lane 0 takes part in every round that any lane does
    for (int k = 0; k < 2; k++) {
      unsigned lo = __builtin_amdgcn_readlane((unsigned)cm[k], 0), hi = __builtin_amdgcn_readlane((unsigned)(cm[k] >> 32), 0);
      cm[k] = ((unsigned long long)hi << 32) | lo;
    }
#endif
    chmask[0] = known ? cm[0] : ~0ull;
    chmask[1] = known ? cm[1] : ~0ull;
  }
  KSYNC();
  if (with_grad) grad_only(m, L, w, nefc);
  return cost;
}

// The cone Hessian of the elliptic contacts in their cone's middle zone as
// weighted virtual rows after the constraint rows (dense rows only), so that
// J^T diag(Dact) J over all the rows includes it. With f_j the scaled
// tangential frictions, T = |(f_j x_j)|, u_j = f_j x_j / T (a unit vector),
// s = sum_j f_j u_j J_j and the contact's rows J_0 (normal) .. J_(dim-1):
//   J_c^T Hc J_c = Dm mu^2 a a^T - Dm cg Y^T (I - u u^T) Y,  a = J_0 - s,
// Y_j = f_j J_j, and I - u u^T = sum_k w_k w_k^T with w_k = e_k - u_k u, so
// its rows are v_0 = J_0 - s (weight Dm mu^2) and v_k = f_k J_k - u_k s
// (weight -Dm cg, positive in the middle zone, where g = N - mu T < 0).
// Returns the row count, or -1 when the virtual rows do not fit in the row
// buffer (the caller then adds the cone terms entry by entry).
KDEV int cone_virtual_rows(const DModel& m, const Layout& L, float* w, int nefc) {
  int nv = m.nv;
  const int* misc = iptr(w, L.misc);
  const int* act = iptr(w, L.efc_act);
  const float *Jaref = w + L.efc_Jaref, *D = w + L.efc_D, *con = w + L.con;
  float *J = w + L.efc_J, *Dact = w + L.efc_Dact;
  int ncon = misc[MISC_NCON], nr = nefc;
  for (int c = 0; c < ncon; c++) {
    const float* cr = con + CON_STRIDE * c;
    int dim = iptr(cr, CON_NROW)[0], r0 = iptr(cr, CON_ROW0)[0];
    if (dim >= 2 && r0 + dim <= nefc && act[r0] == 2) nr += dim;
  }
  if (nr > L.maxefc) return -1;
  nr = nefc;
  float sc = fric_scale(m, L, w);
  for (int c = 0; c < ncon; c++) {
    const float* cr = con + CON_STRIDE * c;
    int dim = iptr(cr, CON_NROW)[0], r0 = iptr(cr, CON_ROW0)[0];
    if (dim < 2 || r0 + dim > nefc || act[r0] != 2) continue;
    const float* f5 = m.pair_friction + 5 * iptr(cr, CON_PAIR)[0];
    const float* x = Jaref + r0;
    float mu = f5[0] * sc, T = 0.f;
    for (int j = 1; j < dim; j++) { float U = (f5[j - 1] * sc) * x[j]; T += U * U; }
    T = sqrtf(T);
    float Dm = D[r0] / (mu * mu * (1.f + mu * mu)), g = mu * x[0] - mu * T, cg = g * mu / T;
    const float* Jc = J + r0 * nv;
    float* V = J + nr * nv;
    for (int i = LANE; i < nv; i += WAVE) {
      float s = 0.f;
      for (int j = 1; j < dim; j++) { float fj = f5[j - 1] * sc; s += fj * (fj * x[j] / T) * Jc[j * nv + i]; }
      V[i] = Jc[i] - s;
      for (int k = 1; k < dim; k++) {
        float fk = f5[k - 1] * sc;
        V[k * nv + i] = fk * Jc[k * nv + i] - (fk * x[k] / T) * s;
      }
    }
    if (LANE == 0) {
      Dact[nr] = Dm * mu * mu;
      for (int k = 1; k < dim; k++) Dact[nr + k] = -Dm * cg;
    }
    nr += dim;
  }
  KSYNC();
  return nr;
}

// H = M + J^T diag(D act) J (FP32 accumulate), factored in FP64 as
// H = L D L^T (right-looking, one LDS phase per column, the column left
// unscaled); the reciprocal pivots go to dvec.
// TRI_CACHE: keep each lane's triangle coordinates in registers across the
// elimination steps for six rounds of lanes (nv <= 27 on 64 lanes); three in
// the register-capped kernel variant (nv <= 20), where more live registers
// cost more spills than they save
template <bool TRI_CACHE>
KDEV int newton_factor(const DModel& m, const Layout& L, float* w, int nefc, KProfAcc* pa) {
  int nv = m.nv;
  double* H = dptr(w, L.H);
  double* idg = dptr(w, L.dvec);
  const float *M = w + L.qM, *J = w + L.efc_J, *Dact = w + L.efc_Dact;
  int ntri = nv * (nv + 1) / 2;
  // rows of J^T diag(Dact) J: the constraint rows and, with dense rows, the
  // elliptic cones' virtual rows
  int nrow = nefc;
  bool cone_rows = false;
  if (m.cone != 0 && !(MJK_BIG && L.jw > 0)) {
    int nr = cone_virtual_rows(m, L, w, nefc);
    if (nr >= 0) { nrow = nr; cone_rows = true; }
  }
#if WAVE == 64
  if (nv <= 32) {
    // J^T diag(Dact) J on the matrix cores: 16 x 16 x 4 FP32 MFMA tiles over the
    // 32 x 32 padding, four constraint rows per step. Operand layout: lane l
    // supplies A[l % 16][l / 16] (= J[r + l / 16][row]) and B[l / 16][l % 16]
    // (= Dact J[r + l / 16][col]) and holds D[4 (l / 16) + q][l % 16].
    typedef float v4f_t __attribute__((ext_vector_type(4)));
    v4f_t t00 = {0.f, 0.f, 0.f, 0.f}, t10 = t00, t11 = t00;
    int lr = LANE & 15, lk = LANE >> 4;
    for (int r0 = 0; r0 < nrow; r0 += 4) {
      int r = r0 + lk;
      bool in = r < nrow;
      float d = in ? Dact[r] : 0.f;
      float a0 = (in && lr < nv) ? J[r * nv + lr] : 0.f;
      float a1 = (in && lr + 16 < nv) ? J[r * nv + 16 + lr] : 0.f;
      float b0 = a0 * d, b1 = a1 * d;
      t00 = __builtin_amdgcn_mfma_f32_16x16x4f32(a0, b0, t00, 0, 0, 0);
      if (nv > 16) {
        t10 = __builtin_amdgcn_mfma_f32_16x16x4f32(a1, b0, t10, 0, 0, 0);
        t11 = __builtin_amdgcn_mfma_f32_16x16x4f32(a1, b1, t11, 0, 0, 0);
      }
    }
    int col = LANE & 15, row0 = 4 * (LANE >> 4);
    for (int q = 0; q < 4; q++) {
      int i = row0 + q;
      if (i < nv && col <= i) H[hx(i, col)] = (double)(M[i * nv + col] + t00[q]);
      if (16 + i < nv && col < nv) H[hx(16 + i, col)] = (double)(M[(16 + i) * nv + col] + t10[q]);
      if (16 + i < nv && 16 + col <= 16 + i) H[hx(16 + i, 16 + col)] = (double)(M[(16 + i) * nv + 16 + col] + t11[q]);
    }
    KSYNC();
  } else
#endif
  if (MJK_BIG && L.jw > 0) {
    // compressed rows: H = M, then each row group sharing its entries' dofs
    // (a friction or limit row, a contact's rows) adds D_r J_r^T J_r into
    // its dofs' entries; the groups go one after the other (they share H
    // entries), the entry pairs of one group in parallel (distinct entries)
    const int jw = L.jw;
    const int* ix = iptr(w, L.efc_Jidx);
    const int* etype = iptr(w, L.efc_type);
    bool compact = m_compact(m);
    for (int t = LANE; t < ntri; t += WAVE) {
      int i, j;
      tri_rc(t, &i, &j);
      H[hx(i, j)] = compact ? 0.0 : (double)M[i * nv + j];
    }
    KSYNC();
    if (compact) {  // the tree entries of the compact rows
      for (int i = LANE; i < nv; i += WAVE) {
        const MJK_G int* ch = m.dof_chain + i * MAX_DOF_DEPTH;
        for (int a = 0, n = m.dof_nchain[i]; a < n; a++) H[hx(i, ch[a])] = (double)M[m.dof_madr[i] + a];
      }
      KSYNC();
    }
    const int* misc = iptr(w, L.misc);
    const float* con = w + L.con;
    int ncon = misc[MISC_NCON], npr = jw * (jw + 1) / 2;
    int nfront = 0;
    while (nfront < nefc && (etype[nfront] == EFC_FRICTION_DOF || etype[nfront] == EFC_LIMIT_JOINT)) nfront++;
    for (int g = 0; g < nfront + ncon; g++) {
      int ra = g, rb = g + 1;
      if (g >= nfront) {
        const float* cr = con + CON_STRIDE * (g - nfront);
        int cnt = iptr(cr, CON_NROW)[0];
        if (cnt == 0) continue;
        ra = iptr(cr, CON_ROW0)[0];
        rb = ra + cnt;
      }
      bool any = false;
      for (int r = ra; r < rb; r++) any = any || Dact[r] != 0.f;
      if (!any) continue;
      for (int t = LANE; t < npr; t += WAVE) {
        int a, b;
        tri_rc(t, &a, &b);
        int ia = ix[ra * jw + a], ib = ix[ra * jw + b];
        if (ia < 0 || ib < 0) continue;
        float sacc = 0.f;
        for (int r = ra; r < rb; r++) sacc += Dact[r] * J[r * jw + a] * J[r * jw + b];
        int i = ia > ib ? ia : ib, j = ia > ib ? ib : ia;
        H[hx(i, j)] += (double)sacc;
      }
      KSYNC();
    }
  } else {
    for (int t = LANE; t < ntri; t += WAVE) {
      int i, j;
      tri_rc(t, &i, &j);
      H[hx(i, j)] = (double)(m_entry(m, M, i, j) + kdot3(Dact, J + i, nv, J + j, nv, nrow));
    }
    KSYNC();
  }
  if (m.cone != 0 && !cone_rows) {
    // contacts in the middle zone of the elliptic cone: H += J_c^T Hc J_c with
    // Hc = Dm (a a^T + g B), a = (mu, -mu f_j u_j), g = N - mu T,
    // B_jk = -mu f_j f_k (delta_jk - u_j u_k) / T (u = U_t / T); per entry
    // J_c^T Hc J_c = Dm [mu^2 (J0i - s_i)(J0l - s_l) - g mu / T (q_il - s_i s_l)],
    // s_i = sum_j f_j^2 x_j J_ji / T, q_il = sum_j f_j^2 J_ji J_jl
    const int* misc = reinterpret_cast<const int*>(w + L.misc);
    const int* act = reinterpret_cast<const int*>(w + L.efc_act);
    const float *Jaref = w + L.efc_Jaref, *D = w + L.efc_D, *con = w + L.con;
    int ncon = misc[MISC_NCON];
    for (int c = 0; c < ncon; c++) {
      const float* cr = con + CON_STRIDE * c;
      int dim = reinterpret_cast<const int*>(cr + CON_NROW)[0], r0 = reinterpret_cast<const int*>(cr + CON_ROW0)[0];
      if (dim < 2 || r0 + dim > nefc || act[r0] != 2) continue;
      const float* f5 = m.pair_friction + 5 * reinterpret_cast<const int*>(cr + CON_PAIR)[0];
      float sc = fric_scale(m, L, w);
      const float* x = Jaref + r0;
      float mu = f5[0] * sc, T = 0.f;
      for (int j = 1; j < dim; j++) { float U = (f5[j - 1] * sc) * x[j]; T += U * U; }
      T = sqrtf(T);
      float Dm = D[r0] / (mu * mu * (1.f + mu * mu)), g = mu * x[0] - mu * T, cg = g * mu / T;
      const float* Jc = J + r0 * nv;
      for (int t = LANE; t < ntri; t += WAVE) {
        int i, l;
        tri_rc(t, &i, &l);
        float si = 0.f, sl = 0.f, q = 0.f;
        for (int j = 1; j < dim; j++) {
          float fj = f5[j - 1] * sc, f2 = fj * fj, ji = Jc[j * nv + i], jl = Jc[j * nv + l];
          si += f2 * x[j] * ji; sl += f2 * x[j] * jl; q += f2 * ji * jl;
        }
        si /= T; sl /= T;
        float hv = Dm * (mu * mu * (Jc[i] - si) * (Jc[l] - sl) - cg * (q - si * sl));
        H[hx(i, l)] += (double)hv;
      }
    }
  }
  KSYNC();
  KPROF_MARK(pa, KP_SPARE);  // profiling builds: the Hessian build is reported apart
  int bad = 0;
  // A lane's entries of a trailing lower triangle, enumerated row-major, have
  // the same (row, col) at every elimination step: for up to TRI_R rounds of
  // lanes they are computed once here (packed row | col << 16) and held in
  // registers by the unrolled loops below, instead of tri_rc per entry and step.
  constexpr int TRI_R = WAVE >= 32 ? (TRI_CACHE ? 6 : 3) * (64 / WAVE) : 1;
  int tri_ab[TRI_R];
  bool tri_cached = false;
  if constexpr (WAVE >= 32) {
    tri_cached = ntri <= TRI_R * WAVE;
#pragma unroll
    for (int r = 0; r < TRI_R; r++) {
      int t = LANE + r * WAVE, a = 0, b = 0;
      if (tri_cached && t < ntri) tri_rc(t, &a, &b);
      tri_ab[r] = a | (b << 16);
    }
  }
  // Two columns per LDS phase. For the pair (k, k+1), with c = H[k+1][k] / d0
  // and d1 = H[k+1][k+1] - c H[k+1][k], the trailing entries take both rank-1
  // updates at once, computing column k+1's updated entries on the fly:
  //   H[i][j] -= H[i][k] H[j][k] / d0 + b_i b_j / d1,  b_x = H[x][k+1] - c H[x][k].
  // Column k+1 itself (its diagonal included) is updated in one pass at the end.
  int k = 0;
  for (; k + 1 < nv - 1; k += 2) {
    double d0 = H[hx(k, k)], h10 = H[hx(k + 1, k)], h11 = H[hx(k + 1, k + 1)];
    if (!(d0 > 0.0)) { bad = 1; d0 = 1e-30; }
    // the pair's second pivot d1 = h11 - h10^2 / d0 = det / d0: 1 / d0 and
    // 1 / det do not wait for each other
    double det = d0 * h11 - h10 * h10;
    if (!(det > 0.0)) { bad = 1; det = 1e-30 * d0; }
    double i0 = krcp(d0), idet = krcp(det);
    double c = h10 * i0, i1 = d0 * idet;
    if (LANE == 0) idg[k] = i0;  // the deferred pass below takes c from it
    int mm = nv - k - 2, nt = mm * (mm + 1) / 2;
    if constexpr (WAVE >= 32) {
      if (tri_cached) {
#pragma unroll
        for (int r = 0; r < TRI_R; r++) {
          if (LANE + r * WAVE >= nt) continue;
          int i = k + 2 + (tri_ab[r] & 0xffff), j = k + 2 + (tri_ab[r] >> 16), ri = hx(i, 0), rj = hx(j, 0);
          double ai = H[ri + k], aj = H[rj + k];
          double bi = H[ri + k + 1] - c * ai, bj = H[rj + k + 1] - c * aj;
          H[ri + j] -= ai * aj * i0 + bi * bj * i1;
        }
        KSYNC();
        continue;
      }
    }
    for (int t = LANE; t < nt; t += WAVE) {
      int a, b;
      tri_rc(t, &a, &b);
      int i = k + 2 + a, j = k + 2 + b, ri = hx(i, 0), rj = hx(j, 0);
      double ai = H[ri + k], aj = H[rj + k];
      double bi = H[ri + k + 1] - c * ai, bj = H[rj + k + 1] - c * aj;
      H[ri + j] -= ai * aj * i0 + bi * bj * i1;
    }
    KSYNC();
  }
  if (k < nv - 1) {  // one column left with trailing rows
    double d0 = H[hx(k, k)];
    if (!(d0 > 0.0)) { bad = 1; d0 = 1e-30; }
    double ik = krcp(d0);
    int nt = (nv - k - 1) * (nv - k) / 2;
    bool done = false;
    if constexpr (WAVE >= 32) {
      if (tri_cached) {
#pragma unroll
        for (int r = 0; r < TRI_R; r++) {
          if (LANE + r * WAVE >= nt) continue;
          int i = k + 1 + (tri_ab[r] & 0xffff), j = k + 1 + (tri_ab[r] >> 16), ri = hx(i, 0);
          H[ri + j] -= H[ri + k] * ik * H[hx(j, k)];
        }
        done = true;
      }
    }
    for (int t = LANE; !done && t < nt; t += WAVE) {
      int a, b;
      tri_rc(t, &a, &b);
      int i = k + 1 + a, j = k + 1 + b, ri = hx(i, 0);
      H[ri + j] -= H[ri + k] * ik * H[hx(j, k)];
    }
    KSYNC();
  }
  // deferred updates of the second column of each pair: H[i][k+1] -= c H[i][k], i >= k+1
  for (int kk = 0; kk + 1 < nv - 1; kk += 2) {
    double c = H[hx(kk + 1, kk)] * idg[kk];
    for (int i = kk + 1 + LANE; i < nv; i += WAVE) H[hx(i, kk + 1)] -= c * H[hx(i, kk)];
  }
  KSYNC();
  for (int i = LANE; i < nv; i += WAVE) {
    double d = H[hx(i, i)];
    if (!(d > 0.0)) { bad = 1; d = 1e-30; }
    idg[i] = krcp(d);
  }
  KSYNC();
  return wori(bad) ? ST_NOT_PD : 0;
}

// Rank-1 changes of the factored Hessian for rows whose activity changed
// (H += D_r J_r^T J_r for a row that became active, -= for one that left):
// the LDL^T update of Gill, Golub, Murray and Saunders (1974, method C1) on
// newton_factor's storage (H_ij = L_ij d_j below the diagonal, H_jj = d_j,
// dvec = 1 / d), one column step per dof with the row's J entries in lanes.
// Returns false (the caller factors anew) when a pivot would lose more than
// ten digits, which a downdate of M + J^T D J should not do.
KDEV bool newton_update(const DModel& m, const Layout& L, float* w, const unsigned long long* chmask) {
  int nv = m.nv;
  double* H = dptr(w, L.H);
  double* idg = dptr(w, L.dvec);
  const float *J = w + L.efc_J, *D = w + L.efc_D;
  const int* act = iptr(w, L.efc_act);
  if (chmask[0] == ~0ull && chmask[1] == ~0ull) return false;  // rows not known
  for (int q = 0; q < 2; q++) {
    unsigned long long bits = chmask[q];
    while (bits) {
      int r = 64 * q + __builtin_ctzll(bits);
      bits &= bits - 1;
      double a = act[r] ? (double)D[r] : -(double)D[r];
#if WAVE == 64
      double z = LANE < nv ? (double)J[r * nv + LANE] : 0.0;
      double d = LANE < nv ? H[hx(LANE, LANE)] : 1.0, id = LANE < nv ? idg[LANE] : 1.0;
      for (int j = 0; j < nv; j++) {
        double p = rl_d(z, j), dj = rl_d(d, j), idj = rl_d(id, j);
        double dn = dj + a * p * p;
        if (!(dn > 1e-10 * dj)) return false;
        double idn = krcp(dn), beta = p * a * idn;
        a *= dj * idn;
        if (LANE == j) { d = dn; id = idn; }
        if (LANE > j && LANE < nv) {
          double* hp = H + hx(LANE, j);
          double lij = *hp * idj;
          z -= p * lij;
          lij += beta * z;
          *hp = lij * dn;
        }
      }
      if (LANE < nv) { H[hx(LANE, LANE)] = d; idg[LANE] = id; }
      KSYNC();
#elif WAVE == 1
      double* z = dptr(w, L.dvec) + nv;
      for (int i = 0; i < nv; i++) z[i] = (double)J[r * nv + i];
      for (int j = 0; j < nv; j++) {
        double p = z[j], dj = H[hx(j, j)], idj = idg[j];
        double dn = dj + a * p * p;
        if (!(dn > 1e-10 * dj)) return false;
        double idn = krcp(dn), beta = p * a * idn;
        a *= dj * idn;
        for (int i = j + 1; i < nv; i++) {
          double* hp = H + hx(i, j);
          double lij = *hp * idj;
          z[i] -= p * lij;
          lij += beta * z[i];
          *hp = lij * dn;
        }
        H[hx(j, j)] = dn;
        idg[j] = idn;
      }
#else
      (void)a; (void)J; (void)H; (void)idg;
      return false;
#endif
    }
  }
  return true;
}

// Mgrad = H^-1 grad with the factor of newton_factor, in FP64; on the GPU the
// vector is held across lanes and each column step is a readlane broadcast.
KDEV void newton_solve(const DModel& m, const Layout& L, float* w) {
  int nv = m.nv;
  const double* H = dptr(w, L.H);
  const double* idg = dptr(w, L.dvec);
  const float* grad = w + L.grad;
  float* Mg = w + L.Mgrad;
#if WAVE >= 32
  if (MJK_BIG && nv > WAVE) {
    double* y = dptr(w, L.dvec) + nv;
    for (int i = LANE; i < nv; i += WAVE) y[i] = grad[i];
    KSYNC();
    for (int k = 0; k < nv - 1; k++) {
      double yk = y[k] * idg[k];
      for (int i = k + 1 + LANE; i < nv; i += WAVE) y[i] -= H[hx(i, k)] * yk;
      KSYNC();
    }
    for (int i = LANE; i < nv; i += WAVE) y[i] *= idg[i];
    KSYNC();
    for (int k = nv - 1; k >= 1; k--) {
      double xk = y[k];
      for (int i = LANE; i < k; i += WAVE) y[i] -= H[hx(k, i)] * idg[i] * xk;
      KSYNC();
    }
    for (int i = LANE; i < nv; i += WAVE) Mg[i] = (float)y[i];
    KSYNC();
    return;
  }
  // each step's factor entry is loaded one step ahead (see solve_ld)
  int ln = LANE < nv ? LANE : 0;
  double v = LANE < nv ? (double)grad[LANE] : 0.0;
  double id = LANE < nv ? idg[LANE] : 1.0;
  // rows of the packed triangle: reads past a row's end (masked lanes) stay
  // inside the triangle for every ln < nv
  const double* Hr = H + hx(ln, 0);
  double hn = Hr[0];
  for (int k = 0; k < nv - 1; k++) {          // L y = g, L_ik = H_ik / H_kk
    double h = hn;
    if (k + 1 < nv - 1) hn = Hr[k + 1];
    double yk = rl_d(v * id, k);  // lane k's final v_k / d_k: one broadcast instead of two
    if (LANE > k && LANE < nv) v -= h * yk;
  }
  v *= id;                                      // D z = y
  hn = nv > 1 ? H[hx(nv - 1, ln)] : 0.0;
  for (int k = nv - 1; k >= 1; k--) {          // L^T x = z
    double h = hn;
    if (k > 1) hn = H[hx(k - 1, ln)];
    double xk = rl_d(v, k);
    if (LANE < k) v -= h * id * xk;
  }
  KSYNC();
  if (LANE < nv) Mg[LANE] = (float)v;
#else
  double* y = dptr(w, L.dvec) + nv;
  for (int i = 0; i < nv; i++) y[i] = grad[i];
  for (int k = 0; k < nv - 1; k++)
    for (int i = k + 1; i < nv; i++) y[i] -= H[hx(i, k)] * idg[k] * y[k];
  for (int i = 0; i < nv; i++) y[i] *= idg[i];
  for (int k = nv - 1; k >= 1; k--)
    for (int i = 0; i < k; i++) y[i] -= H[hx(k, i)] * idg[i] * y[k];
  for (int i = 0; i < nv; i++) Mg[i] = (float)y[i];
#endif
  KSYNC();
}

// d1n: sum of the magnitudes of the slope's terms, the scale of its FP32
// rounding
struct LsPt { float alpha, f, d1, d2, d1n; int nchange; };

// One row's line-search operands (Jaref, Jv, D, type bits), one 16-byte entry.
struct alignas(16) LsRow4 { float jar, jv, d, ty; };

// Free LDS beside a device-memory workspace: up to cap rows of the line
// search are staged there (the kernels whose workspace is in device memory
// hold their occupancy by registers, so a few KB of LDS per workgroup are free).
struct LsCache { LsRow4* p; int cap; };

// The rows a lane owns (r = LANE + q WAVE) held in registers for the whole
// line search, when there are at most two per lane: every evaluation is then
// register arithmetic plus the wave reductions, with no LDS traffic. More
// rows: staged once in the LDS cache when there is one (every evaluation then
// reads one 16-byte LDS entry per row instead of four workspace words).
struct LsRows { float jar[2], jv[2], d[2]; int ty[2]; int on; const LsRow4* staged; };

KDEV void ls_rows(const Layout& L, const float* w, int nefc, LsRows* c, LsCache lc) {
  c->on = 0;
  c->staged = nullptr;
  if (nefc > 2 * WAVE && nefc <= lc.cap) {
    const float *Jaref = w + L.efc_Jaref, *Jv = w + L.efc_Jv, *D = w + L.efc_D;
    const int* etype = reinterpret_cast<const int*>(w + L.efc_type);
    for (int r = LANE; r < nefc; r += WAVE) {
      LsRow4 e;
      e.jar = Jaref[r]; e.jv = Jv[r]; e.d = D[r]; e.ty = ki2f(etype[r]);
      lc.p[r] = e;
    }
    KSYNC();
    c->staged = lc.p;
    return;
  }
#if WAVE >= 32
  if (nefc <= 2 * WAVE) {
    c->on = 1;
    const float *Jaref = w + L.efc_Jaref, *Jv = w + L.efc_Jv, *D = w + L.efc_D;
    const int* etype = reinterpret_cast<const int*>(w + L.efc_type);
    for (int q = 0; q < 2; q++) {
      int r = LANE + q * WAVE;
      bool in = r < nefc;
      c->jar[q] = in ? Jaref[r] : 0.f;
      c->jv[q] = in ? Jv[r] : 0.f;
      c->d[q] = in ? D[r] : 0.f;
      c->ty[q] = in ? etype[r] : -1;
    }
  }
#else
  (void)L; (void)w; (void)nefc;
#endif
}

// cost terms of one row at alpha (x = Jaref + alpha Jv); zone change against ref_alpha
KDEV void ls_row(const DModel& m, const Layout& L, const float* w, int r, int ty, float jar, float jv, float d, float alpha,
                 float ref_alpha, float& f, float& d1, float& d2, int& nch) {
  if (ty == EFC_CONTACT_ELLIPTIC) return;  // evaluated per contact
  float x = jar + alpha * jv, xr = jar + ref_alpha * jv;
  if (ty == EFC_FRICTION_DOF) {
    float fl = m.dof_frictionloss[iptr(w, L.efc_id)[r]], R = 1.f / d;
    int z = x <= -R * fl ? 0 : (x >= R * fl ? 2 : 1);
    int zr = xr <= -R * fl ? 0 : (xr >= R * fl ? 2 : 1);
    nch += z != zr;
    if (z == 0) { f += -fl * x - 0.5f * R * fl * fl; d1 += -fl * jv; }
    else if (z == 2) { f += fl * x - 0.5f * R * fl * fl; d1 += fl * jv; }
    else { f += 0.5f * d * x * x; d1 += d * x * jv; d2 += d * jv * jv; }
  } else if (ty == EFC_EQUALITY) {
    f += 0.5f * d * x * x; d1 += d * x * jv; d2 += d * jv * jv;
  } else {
    nch += (x < 0.f) != (xr < 0.f);
    if (x < 0.f) { f += 0.5f * d * x * x; d1 += d * x * jv; d2 += d * jv * jv; }
  }
}

// cost along search at alpha; nchange counts rows whose zone differs from
// their zone at ref_alpha (0 means one quadratic piece spans both points)
KDEV LsPt ls_eval(const DModel& m, const Layout& L, float* w, int nefc, float alpha, float ref_alpha, const float* gq,
                  const LsRows& c) {
  float f = 0.f, d1 = 0.f, d2 = 0.f, d1a = 0.f;
  int nch = 0;
  if (m.cone != 0) {
    // elliptic contacts; the middle zone is not quadratic, so it never counts as one exact piece
    const int* misc = reinterpret_cast<const int*>(w + L.misc);
    int ncon = misc[MISC_NCON];
    const float *Jaref = w + L.efc_Jaref, *Jv = w + L.efc_Jv, *D = w + L.efc_D, *con = w + L.con;
    for (int q = LANE; q < ncon; q += WAVE) {
      const float* cr = con + CON_STRIDE * q;
      int dim = reinterpret_cast<const int*>(cr + CON_NROW)[0], r0 = reinterpret_cast<const int*>(cr + CON_ROW0)[0];
      if (dim < 2 || r0 + dim > nefc) continue;
      const float* f5 = m.pair_friction + 5 * reinterpret_cast<const int*>(cr + CON_PAIR)[0];
      float sc = fric_scale(m, L, w), cv, g1, g2, cr_, h1, h2;
      int z = k_ellip_line(Jaref + r0, Jv + r0, D + r0, f5, sc, dim, alpha, &cv, &g1, &g2);
      int zr = k_ellip_line(Jaref + r0, Jv + r0, D + r0, f5, sc, dim, ref_alpha, &cr_, &h1, &h2);
      nch += (z != zr) || z == 2;
      f += cv; d1 += g1; d2 += g2; d1a += fabsf(g1);
    }
  }
  if (c.on) {
    for (int q = 0; q < 2; q++)
      if (c.ty[q] >= 0) {
        float t = d1;
        ls_row(m, L, w, LANE + q * WAVE, c.ty[q], c.jar[q], c.jv[q], c.d[q], alpha, ref_alpha, f, d1, d2, nch);
        d1a += fabsf(d1 - t);
      }
  } else if (c.staged) {
    for (int r = LANE; r < nefc; r += WAVE) {
      LsRow4 e = c.staged[r];
      float t = d1;
      ls_row(m, L, w, r, kf2i(e.ty), e.jar, e.jv, e.d, alpha, ref_alpha, f, d1, d2, nch);
      d1a += fabsf(d1 - t);
    }
  } else {
    const float *Jaref = w + L.efc_Jaref, *Jv = w + L.efc_Jv, *D = w + L.efc_D;
    const int* etype = iptr(w, L.efc_type);
    for (int r = LANE; r < nefc; r += WAVE) {
      float t = d1;
      ls_row(m, L, w, r, etype[r], Jaref[r], Jv[r], D[r], alpha, ref_alpha, f, d1, d2, nch);
      d1a += fabsf(d1 - t);
    }
  }
  LsPt p;
  p.alpha = alpha;
  p.f = wsum(f) + gq[0] + alpha * gq[1] + 0.5f * alpha * alpha * gq[2];
  p.d1 = wsum(d1) + gq[1] + alpha * gq[2];
  p.d1n = wsum(d1a) + fabsf(gq[1]) + fabsf(alpha * gq[2]);
  p.d2 = wsum(d2) + gq[2];
  p.nchange = wsumi(nch);
  return p;
}

// the slope is zero to the tolerance, or to FP32 resolution (a few 1e-7 of
// its terms' magnitudes, below which further steps only chase rounding)
KDEV bool ls_flat(const LsPt& p, float gtol) { return fabsf(p.d1) <= fmaxf(gtol, 1e-6f * p.d1n); }

// Decrease of the constraint problem's cost from alpha = 0 to alpha along the
// search direction, summed from per-row changes (a quadratic row active at
// both points changes by 1/2 D alpha Jv (x0 + x)) and the Gauss term's
// alpha g1 + 1/2 alpha^2 g2: MuJoCo's "improvement" (old cost - new cost)
// without the cancellation of two large FP32 sums, which near the optimum is
// the size of the improvement itself.
KDEV float ls_decrease(const DModel& m, const Layout& L, const float* w, int nefc, float alpha, const float* gq,
                       const LsRow4* staged) {
  const float *Jaref = w + L.efc_Jaref, *Jv = w + L.efc_Jv, *D = w + L.efc_D;
  const int* etype = iptr(w, L.efc_type);
  float df = 0.f;
  for (int r = LANE; r < nefc; r += WAVE) {
    LsRow4 e;
    if (staged) e = staged[r];
    else { e.jar = Jaref[r]; e.jv = Jv[r]; e.d = D[r]; e.ty = ki2f(etype[r]); }
    int ty = kf2i(e.ty);
    if (ty == EFC_CONTACT_ELLIPTIC) continue;
    float x0 = e.jar, jv = e.jv, d = e.d, x = x0 + alpha * jv;
    if (ty == EFC_FRICTION_DOF) {
      float fl = m.dof_frictionloss[iptr(w, L.efc_id)[r]], R = 1.f / d;
      int z = x <= -R * fl ? 0 : (x >= R * fl ? 2 : 1);
      int z0 = x0 <= -R * fl ? 0 : (x0 >= R * fl ? 2 : 1);
      if (z == z0) df += z == 1 ? 0.5f * d * alpha * jv * (x0 + x) : (z == 0 ? -fl : fl) * alpha * jv;
      else {
        float c = z == 1 ? 0.5f * d * x * x : (z == 0 ? -fl * x : fl * x) - 0.5f * R * fl * fl;
        float c0 = z0 == 1 ? 0.5f * d * x0 * x0 : (z0 == 0 ? -fl * x0 : fl * x0) - 0.5f * R * fl * fl;
        df += c - c0;
      }
    } else if (ty == EFC_EQUALITY || (x < 0.f && x0 < 0.f)) {
      df += 0.5f * d * alpha * jv * (x0 + x);
    } else if (x < 0.f) {
      df += 0.5f * d * x * x;
    } else if (x0 < 0.f) {
      df -= 0.5f * d * x0 * x0;
    }
  }
  if (m.cone != 0) {
    const int* misc = reinterpret_cast<const int*>(w + L.misc);
    int ncon = misc[MISC_NCON];
    const float* con = w + L.con;
    for (int q = LANE; q < ncon; q += WAVE) {
      const float* cr = con + CON_STRIDE * q;
      int dim = reinterpret_cast<const int*>(cr + CON_NROW)[0], r0 = reinterpret_cast<const int*>(cr + CON_ROW0)[0];
      if (dim < 2 || r0 + dim > nefc) continue;
      const float* f5 = m.pair_friction + 5 * reinterpret_cast<const int*>(cr + CON_PAIR)[0];
      float sc = fric_scale(m, L, w), c, c0, g1, g2;
      k_ellip_line(Jaref + r0, Jv + r0, D + r0, f5, sc, dim, alpha, &c, &g1, &g2);
      k_ellip_line(Jaref + r0, Jv + r0, D + r0, f5, sc, dim, 0.f, &c0, &g1, &g2);
      df += c - c0;
    }
  }
  return -(wsum(df) + alpha * gq[1] + 0.5f * alpha * alpha * gq[2]);
}

// Exact line search on the convex piecewise-quadratic cost (same algorithm as
// the oracle): Newton steps until the derivative changes sign, then Newton
// from the left end / secant on the bracket. A Newton step whose end point is
// in the same quadratic piece as its start is the exact minimiser.
KDEV float linesearch_(const DModel& m, const Layout& L, float* w, int nefc, int* exact, float* gq_out, LsCache lc,
                       const LsRow4** staged, const float* newton_cost);

// dec (when not null): the cost decrease at the returned step, by ls_decrease.
// newton_cost (when not null): the search direction is the Newton direction
// of the current point, whose cost it holds
KDEV float linesearch(const DModel& m, const Layout& L, float* w, int nefc, int* exact, float* dec, LsCache lc,
                      const float* newton_cost) {
  float gq[3];
  const LsRow4* staged = nullptr;
  float a = linesearch_(m, L, w, nefc, exact, gq, lc, &staged, newton_cost);
  if (dec) *dec = a > 0.f ? ls_decrease(m, L, w, nefc, a, gq, staged) : 0.f;
  return a;
}

KDEV float linesearch_(const DModel& m, const Layout& L, float* w, int nefc, int* exact, float* gq_out, LsCache lc,
                       const LsRow4** staged, const float* newton_cost) {
  int nv = m.nv;
  const float *s = w + L.search, *M = w + L.qM;
  float *Mv = w + L.Mv, *Jv = w + L.efc_Jv;
  const float *Ma = w + L.Ma, *qfs = w + L.qfrc_smooth, *qacc = w + L.qacc, *qas = w + L.qacc_smooth;
  float p0s = 0.f;
  for (int i = LANE; i < nv; i += WAVE) p0s += s[i] * s[i];
  float snorm = sqrtf(wsum(p0s));
  if (!(snorm >= KMINVAL)) return 0.f;
  float scale = 1.f / (m.meaninertia * (float)(nv > 1 ? nv : 1));
  float gtol = m.tolerance * m.ls_tolerance * snorm / scale;
  // compressed rows in device memory: J s gathers s from a copy in the line
  // search's LDS cache, which the rows are staged into only afterwards
  const float* sx = s;
  if (MJK_BIG && L.jw > 0 && lc.p && 4 * lc.cap >= nv) {
    float* xs = reinterpret_cast<float*>(lc.p);
    for (int i = LANE; i < nv; i += WAVE) xs[i] = s[i];
    KSYNC();
    sx = xs;
  }
  mulM(m, M, s, Mv);
  jrows_dot(m, L, w, nefc, sx, Jv, nullptr);
  KSYNC();
  float g0 = 0.f, g1 = 0.f, g2 = 0.f;
  for (int i = LANE; i < nv; i += WAVE) {
    g0 += 0.5f * (Ma[i] - qfs[i]) * (qacc[i] - qas[i]);
    g1 += (Ma[i] - qfs[i]) * s[i];
    g2 += s[i] * Mv[i];
  }
  float gq[3] = {wsum(g0), wsum(g1), wsum(g2)};
  gq_out[0] = gq[0]; gq_out[1] = gq[1]; gq_out[2] = gq[2];
  *exact = 0;
  LsRows rows;
  ls_rows(L, w, nefc, &rows, lc);
  *staged = rows.staged;
  LsPt p0;
  if (newton_cost) {
    // s = -H^-1 g: the slope along s at alpha = 0 is g.s and the curvature
    // s'Hs = -g.s (H the Hessian s was solved with, that of the current
    // active set), so the start point needs no pass over the rows
    const float* g = w + L.grad;
    float gs = 0.f;
    for (int i = LANE; i < nv; i += WAVE) gs += g[i] * s[i];
    gs = wsum(gs);
    p0.alpha = 0.f; p0.f = *newton_cost; p0.d1 = gs; p0.d2 = -gs; p0.d1n = fabsf(gs); p0.nchange = 0;
  } else {
    p0 = ls_eval(m, L, w, nefc, 0.f, 0.f, gq, rows);
  }
  if (!(p0.d1 < 0.f)) return 0.f;
  int it = 1, maxit = m.ls_iterations;
  LsPt lo = p0;
  LsPt pt = ls_eval(m, L, w, nefc, -p0.d1 / p0.d2, 0.f, gq, rows);
  if (pt.nchange == 0) { *exact = 1; return pt.alpha; }
  if (ls_flat(pt, gtol)) return pt.alpha;
  while (pt.d1 < 0.f && it < maxit) {
    lo = pt;
    pt = ls_eval(m, L, w, nefc, pt.alpha - pt.d1 / pt.d2, lo.alpha, gq, rows);
    it++;
    if (pt.nchange == 0 || ls_flat(pt, gtol)) return pt.alpha;
  }
  if (pt.d1 < 0.f) return pt.alpha;
  LsPt hi = pt;
  while (it < maxit) {
    float a = lo.alpha - lo.d1 / lo.d2;
    bool newton = a > lo.alpha && a < hi.alpha;
    if (!newton) a = lo.alpha - lo.d1 * (hi.alpha - lo.alpha) / (hi.d1 - lo.d1);
    if (!(a > lo.alpha && a < hi.alpha)) a = 0.5f * (lo.alpha + hi.alpha);
    pt = ls_eval(m, L, w, nefc, a, lo.alpha, gq, rows);
    it++;
    if ((newton && pt.nchange == 0) || ls_flat(pt, gtol)) return pt.alpha;
    if (pt.d1 < 0.f) lo = pt; else hi = pt;
    if (!(hi.alpha > lo.alpha)) break;
  }
  return lo.f <= hi.f ? lo.alpha : hi.alpha;
}

// mj_fwdConstraint: warm start choice, Newton or CG iterations with MuJoCo's
// termination test, qfrc_constraint = J^T force.
// pred: the learned warm-start candidate (the env's row of the prediction in
// device memory), or null
template <bool TRI_CACHE>
KDEV int st_solve(const DModel& m, const Layout& L, float* w, int* niter_out, KProfAcc* pa, const float* pred,
                  LsCache lc) {
  int* misc = iptr(w, L.misc);
  int nv = m.nv;
  int nefc = misc[MISC_NEFC];
  float *qacc = w + L.qacc, *qfc = w + L.qfrc_constraint;
  int st = 0;
  *niter_out = 0;
  if (nefc == 0) {
    const float* qas = w + L.qacc_smooth;
    for (int i = LANE; i < nv; i += WAVE) { qacc[i] = qas[i]; qfc[i] = 0.f; }
    KSYNC();
    return 0;
  }
  int nch = 0;
  // warm start: the candidates compared by cost alone (M qacc_smooth =
  // qfrc_smooth needs no product), the gradient only at the one kept
  const float *qas = w + L.qacc_smooth, *qws = w + L.qacc_ws, *qfs0 = w + L.qfrc_smooth;
  bool ws = !(m.disableflags & DS_WARMSTART);
  set_qacc(m, L, w, qas, nefc, qfs0);
  float cost = cost_grad(m, L, w, nefc, &nch, !ws);
  if (ws) {
    float cs = cost;
    set_qacc(m, L, w, qws, nefc);
    cost = cost_grad(m, L, w, nefc, &nch, false);
    const float* best = cost > cs ? qas : qws;
    float cbest = cost > cs ? cs : cost;
    const float* cur = qws;  // the candidate the workspace holds
    if (pred) {
      // learned candidate: kept only when its cost is strictly the lowest
      set_qacc(m, L, w, pred, nefc);
      float cp = cost_grad(m, L, w, nefc, &nch, false);
      if (cp < cbest) { best = pred; cbest = cp; }
      cost = cp;
      cur = pred;
    }
    if (best != cur) {
      set_qacc(m, L, w, best, nefc, best == qas ? qfs0 : nullptr);
      cost = cost_grad(m, L, w, nefc, &nch);
    } else {
      grad_only(m, L, w, nefc);
    }
  }
  float scale = 1.f / (m.meaninertia * (float)(nv > 1 ? nv : 1));
  bool newton = m.solver == 2;
  float *grad = w + L.grad, *Mg = w + L.Mgrad, *srch = w + L.search, *Ma = w + L.Ma, *Mv = w + L.Mv;
  float *go = w + L.grad_old, *mgo = w + L.Mgrad_old, *Jaref = w + L.efc_Jaref;
  const float* Jv = w + L.efc_Jv;
  KPROF_MARK(pa, KP_SOLVE_INIT);
  if (newton) {
    st |= newton_factor<TRI_CACHE>(m, L, w, nefc, pa);
    KPROF_MARK(pa, KP_DIR);
    newton_solve(m, L, w);
    KPROF_MARK(pa, KP_NSOLVE);
  } else {
    for (int i = LANE; i < nv; i += WAVE) Mg[i] = grad[i];
    KSYNC();
    solve_ld(m, w + L.qLD, Mg);
  }
  for (int i = LANE; i < nv; i += WAVE) srch[i] = -Mg[i];
  KSYNC();
  int it;
  for (it = 0; it < m.iterations; it++) {
    int exact = 0;
    KPROF_MARK(pa, KP_DIR);
    float dec = 0.f;
    float alpha = linesearch(kfresh(m), kfresh(L), w, nefc, &exact, newton ? nullptr : &dec, lc,
                             newton ? &cost : nullptr);
    KPROF_MARK(pa, KP_LS);
    *niter_out = it + 1;
    if (alpha == 0.f) break;
    for (int i = LANE; i < nv; i += WAVE) {
      qacc[i] += alpha * srch[i];
      Ma[i] += alpha * Mv[i];
      go[i] = grad[i];
      mgo[i] = Mg[i];
    }
    for (int r = LANE; r < nefc; r += WAVE) Jaref[r] += alpha * Jv[r];
    KSYNC();
    float oldcost = cost;
    KPROF_MARK(pa, KP_SOLVE_REST);
    unsigned long long chm[2];
    cost = cost_grad(kfresh(m), kfresh(L), w, nefc, &nch, true, chm);
    KPROF_MARK(pa, KP_COSTGRAD);
    float gp = 0.f;
    for (int i = LANE; i < nv; i += WAVE) gp += grad[i] * grad[i];
    // CG: the decrease summed per row (ls_decrease); Newton ends on an exact
    // quadratic piece or on the difference of the two costs
    float improvement = scale * (newton ? oldcost - cost : dec), gradient = scale * sqrtf(wsum(gp));
    if (improvement < m.tolerance || gradient < m.tolerance) break;
    // a Newton step that stayed inside one quadratic piece (no row changed
    // zone) landed on that piece's stationary point: the exact minimiser
    if (newton && exact) break;
    if (newton) {
      // H depends on qacc only through the active set: refactor only when it changed
      // one or two rows changed activity: rank-1 updates of the factor
      // (dense rows only: the compressed-row Hessian is factored anew)
      if (nch && !(nch <= 2 && (!MJK_BIG || L.jw == 0) && newton_update(kfresh(m), kfresh(L), w, chm)))
        st |= newton_factor<TRI_CACHE>(kfresh(m), kfresh(L), w, nefc, pa);
      KPROF_MARK(pa, KP_DIR);
      newton_solve(kfresh(m), kfresh(L), w);
      KPROF_MARK(pa, KP_NSOLVE);
    } else {
      for (int i = LANE; i < nv; i += WAVE) Mg[i] = grad[i];
      KSYNC();
      solve_ld(m, w + L.qLD, Mg);
    }
    if (newton) {
      for (int i = LANE; i < nv; i += WAVE) srch[i] = -Mg[i];
    } else {
      float nump = 0.f, denp = 0.f;
      for (int i = LANE; i < nv; i += WAVE) { nump += grad[i] * (Mg[i] - mgo[i]); denp += go[i] * mgo[i]; }
      float num = wsum(nump), den = wsum(denp);
      float beta = num / fmaxf(KMINVAL, den);
      if (beta < 0.f) beta = 0.f;
      for (int i = LANE; i < nv; i += WAVE) srch[i] = -Mg[i] + beta * srch[i];
    }
    KSYNC();
  }
  KPROF_MARK(pa, KP_DIR);
  // qfrc_constraint = J^T force at the final point, which the last
  // cost_grad left in grad = Ma - qfrc_smooth - J^T force
  {
    const float* qfs = w + L.qfrc_smooth;
    for (int i = LANE; i < nv; i += WAVE) qfc[i] = (Ma[i] - qfs[i]) - grad[i];
  }
  KSYNC();
  KPROF_MARK(pa, KP_SOLVE_REST);
  return st;
}

// ------------------------------------------------------------------ integration
// Euler (implicit dof damping when present) / implicitfast, then the position
// update; free and ball quaternions are integrated in FP64
// (reference euler.hpp:26-110 restated with MuJoCo's mj_integratePos).
template <bool DENSE>
KDEV int st_integrate(const DModel& m, const Layout& L, float* w) {
  int nv = m.nv;
  float *qvel = w + L.qvel, *qpos = w + L.qpos, *qacc = w + L.qacc, *ws = w + L.qacc_ws;
  float* acc = w + L.tmpv;
  float h = m.timestep;
  int st = 0;
  bool implicit = m.integrator == 2 || m.integrator == 3;
  bool damp = m.has_damping && !(m.disableflags & DS_EULERDAMP);
  if (implicit || damp) {
    float* A = w + L.qLD;  // scratch: qLD of M is dead after the solver
    float* Bm = w + L.qM;   // M + h*D built in place of qM (dead after the solver)
    for (int k = LANE; k < nv; k += WAVE) Bm[mxi(m, k, k)] += h * m.dof_damping[k];
    KSYNC();
    if (implicit) {
      const float* af = w + L.act_force;
      for (int k = LANE; k < nv; k += WAVE) {
        float add = 0.f;
        for (int t = m.dof_act_adr[k]; t < m.dof_act_adr[k] + m.dof_act_num[k]; t++) {
          int u = m.dof_act_list[t];
          if (m.disableflags & DS_ACTUATION) break;
          float f = af[u];
          if (m.act_forcelimited[u] && (f <= m.act_forcerange[2 * u] || f >= m.act_forcerange[2 * u + 1])) continue;
          add -= h * m.act_gear[u] * m.act_gear[u] * m.act_bias[3 * u + 2];
        }
        Bm[mxi(m, k, k)] += add;
      }
      KSYNC();
    }
    st |= factor_ld<DENSE>(m, Bm, A);
    const float *qfs = w + L.qfrc_smooth, *qfc = w + L.qfrc_constraint;
    for (int k = LANE; k < nv; k += WAVE) acc[k] = qfs[k] + qfc[k];
    KSYNC();
    solve_ld(m, A, acc);
  } else {
    for (int k = LANE; k < nv; k += WAVE) acc[k] = qacc[k];
    KSYNC();
  }
  for (int k = LANE; k < nv; k += WAVE) qvel[k] += h * acc[k];
  KSYNC();
  for (int j = LANE; j < m.njnt; j += WAVE) {
    int qa = m.jnt_qposadr[j], da = m.jnt_dofadr[j], jt = m.jnt_type[j];
    if (jt == JFREE) {
      for (int i = 0; i < 3; i++) qpos[qa + i] += h * qvel[da + i];
      qa += 3; da += 3;
    }
    if (jt == JFREE || jt == JBALL) {
      double wv[3] = {qvel[da], qvel[da + 1], qvel[da + 2]};
      double wn = sqrt(wv[0] * wv[0] + wv[1] * wv[1] + wv[2] * wv[2]);
      double q[4] = {qpos[qa], qpos[qa + 1], qpos[qa + 2], qpos[qa + 3]};
      double qn = sqrt(q[0] * q[0] + q[1] * q[1] + q[2] * q[2] + q[3] * q[3]);
      if (qn < 1e-15) { q[0] = 1; q[1] = q[2] = q[3] = 0; }
      else for (int i = 0; i < 4; i++) q[i] /= qn;
      double qr[4] = {1, 0, 0, 0};
      if (wn >= 1e-15) {
        double ang = 0.5 * (double)h * wn, sa = sin(ang) / wn;
        qr[0] = cos(ang); qr[1] = wv[0] * sa; qr[2] = wv[1] * sa; qr[3] = wv[2] * sa;
      }
      double o0 = q[0] * qr[0] - q[1] * qr[1] - q[2] * qr[2] - q[3] * qr[3];
      double o1 = q[0] * qr[1] + q[1] * qr[0] + q[2] * qr[3] - q[3] * qr[2];
      double o2 = q[0] * qr[2] - q[1] * qr[3] + q[2] * qr[0] + q[3] * qr[1];
      double o3 = q[0] * qr[3] + q[1] * qr[2] - q[2] * qr[1] + q[3] * qr[0];
      double on = sqrt(o0 * o0 + o1 * o1 + o2 * o2 + o3 * o3);
      qpos[qa] = (float)(o0 / on); qpos[qa + 1] = (float)(o1 / on);
      qpos[qa + 2] = (float)(o2 / on); qpos[qa + 3] = (float)(o3 / on);
    } else {
      qpos[qa] += h * qvel[da];
    }
  }
  for (int k = LANE; k < nv; k += WAVE) ws[k] = qacc[k];
  KSYNC();
  return st;
}

// ------------------------------------------------------------------ one env
// aux: optional per-env outputs of the last substep's pre-integration state
// (xpos, xquat, subtree_com, cinert, cvel, qfrc_actuator).
struct AuxOut {
  float *xpos, *xquat, *subtree_com, *cinert, *cvel, *qfrc_actuator, *cfrc_ext;
};

// Global arrays of one launch. State arrays are env-major (env * dim + i), as
// the C API exchanges them (reference include/mjmlx/mjmlx.h:137-138,206-216).
// ------------------------------------------------------------------ rays
// mj_ray against one primitive in its own frame: the smallest x >= 0 with
// p + x v on the surface, or -1 (v need not be unit length; x is in units of
// |v|). Planes are hit from their +z side only, within their half extents
// when those are nonzero; a ray starting inside a closed primitive hits its
// exit point. Reference src/batched.cpp:1739-1806.
KDEV float ray_quadric(float a, float b, float c) {  // smallest x >= 0 of a x^2 + 2 b x + c
  float disc = b * b - a * c;
  if (disc < 1e-12f || fabsf(a) < 1e-12f) return -1.f;
  float s = sqrtf(disc), x0 = (-b - s) / a, x1 = (-b + s) / a;
  return x0 >= 0.f ? x0 : (x1 >= 0.f ? x1 : -1.f);
}
KDEV float ray_closer(float best, float x) { return x >= 0.f && (best < 0.f || x < best) ? x : best; }
KDEV float ray_geom(int type, const float* sz, const float* p, const float* v) {
  switch (type) {
    case GPLANE: {
      if (v[2] > -1e-9f) return -1.f;
      float x = -p[2] / v[2];
      if (x < 0.f) return -1.f;
      float hx = p[0] + x * v[0], hy = p[1] + x * v[1];
      if ((sz[0] > 0.f && fabsf(hx) > sz[0]) || (sz[1] > 0.f && fabsf(hy) > sz[1])) return -1.f;
      return x;
    }
    case GSPHERE:
      return ray_quadric(dot3(v, v), dot3(p, v), dot3(p, p) - sz[0] * sz[0]);
    case GELLIPSOID: {
      float ps[3] = {p[0] / sz[0], p[1] / sz[1], p[2] / sz[2]}, vs[3] = {v[0] / sz[0], v[1] / sz[1], v[2] / sz[2]};
      return ray_quadric(dot3(vs, vs), dot3(ps, vs), dot3(ps, ps) - 1.f);
    }
    case GCAPSULE:
    case GCYLINDER: {
      float r = sz[0], h = sz[1], best = -1.f;
      // lateral surface |(x, y)| = r within |z| <= h
      float a = v[0] * v[0] + v[1] * v[1], b = p[0] * v[0] + p[1] * v[1], c = p[0] * p[0] + p[1] * p[1] - r * r;
      float disc = b * b - a * c;
      if (a > 1e-12f && disc >= 0.f) {
        float sd = sqrtf(disc);
        for (int k = -1; k <= 1; k += 2) {
          float x = (-b + k * sd) / a;
          if (fabsf(p[2] + x * v[2]) <= h) best = ray_closer(best, x);
        }
      }
      for (int e = -1; e <= 1; e += 2) {
        if (type == GCYLINDER) {  // flat end discs
          if (fabsf(v[2]) <= 1e-12f) continue;
          float x = (e * h - p[2]) / v[2];
          float qx = p[0] + x * v[0], qy = p[1] + x * v[1];
          if (x >= 0.f && qx * qx + qy * qy <= r * r) best = ray_closer(best, x);
        } else {  // hemispherical ends: the outer half of the sphere at z = e h
          float pc[3] = {p[0], p[1], p[2] - e * h};
          float a2 = dot3(v, v), b2 = dot3(pc, v), c2 = dot3(pc, pc) - r * r, d2 = b2 * b2 - a2 * c2;
          if (d2 < 0.f) continue;
          float sd = sqrtf(d2);
          for (int k = -1; k <= 1; k += 2) {
            float x = (-b2 + k * sd) / a2;
            if (e * (p[2] + x * v[2] - e * h) >= 0.f) best = ray_closer(best, x);
          }
        }
      }
      return best;
    }
    case GBOX: {  // slabs
      float tn = -1e30f, tf = 1e30f;
      for (int i = 0; i < 3; i++) {
        if (fabsf(v[i]) < 1e-12f) {
          if (fabsf(p[i]) > sz[i]) return -1.f;
          continue;
        }
        float t0 = (-sz[i] - p[i]) / v[i], t1 = (sz[i] - p[i]) / v[i];
        tn = fmaxf(tn, fminf(t0, t1));
        tf = fminf(tf, fmaxf(t0, t1));
      }
      if (tn > tf || tf < 0.f) return -1.f;
      return tn >= 0.f ? tn : tf;
    }
  }
  return -1.f;
}

// Rays of one env against its geoms at the current qpos: the env's kinematics
// into the workspace, then one lane per ray over the geoms allowed by gmask.
// rays [env][nray][pnt 3, vec 3]; dist / gid [env][nray] (-1: no hit).
KDEV void run_rays(const DModel& m, const Layout& L, float* w, const float* qpos, const float* rand, int nrand,
                   const float* rays, int nray, const int* gmask, float* dist, int* gid, int env) {
  for (int i = LANE; i < m.nq; i += WAVE) w[L.qpos + i] = qpos[(size_t)env * m.nq + i];
  if (L.rand >= 0)
    for (int i = LANE; i < nrand; i += WAVE) w[L.rand + i] = rand[(size_t)env * nrand + i];
  KSYNC();
  st_kinematics(m, L, w);
  const float *gx = w + L.gxpos, *gm = w + L.gxmat;
  for (int r = LANE; r < nray; r += WAVE) {
    const float* ry = rays + ((size_t)env * nray + r) * 6;
    float best = -1.f;
    int bg = -1;
    for (int g = 0; g < m.ngeom; g++) {
      if (!gmask[g]) continue;
      const float *x = gx + 3 * g, *R = gm + 9 * g;
      float d[3] = {ry[0] - x[0], ry[1] - x[1], ry[2] - x[2]}, lp[3], lv[3];
      for (int k = 0; k < 3; k++) {
        lp[k] = R[k] * d[0] + R[3 + k] * d[1] + R[6 + k] * d[2];
        lv[k] = R[k] * ry[3] + R[3 + k] * ry[4] + R[6 + k] * ry[5];
      }
      float t = ray_geom(m.geom_type[g], m.geom_size + 3 * g, lp, lv);
      if (t >= 0.f && (best < 0.f || t < best)) { best = t; bg = g; }
    }
    dist[(size_t)env * nray + r] = best;
    gid[(size_t)env * nray + r] = bg;
  }
}

struct StepIO {
  float *qpos, *qvel, *qacc_ws;
  const float* ctrl;
  int *niter, *status, *ncon, *nefc;
  int accumulate;            // add niter / OR status into the outputs instead of overwriting them
  const float* qacc_pred;    // learned warm start: predicted qacc [env * nv], or null
  float* feat;               // learned warm start: raw features after the last substep [env * nfeat(+nv)], or null
  int feat_prev_qacc;        // append the substep's converged qacc to the features
  float *dump1, *dump2;  // debug: workspace after constraint build / after solve (last substep)
  AuxOut aux;
  int has_aux;
  unsigned long long* prof;  // profiling builds: per-env stage cycles [env * KP_N + stage]
  const float* rand;         // domain randomisation: per-env parameter blocks [env * nrand], or null
  int nrand;
  float* sensordata;         // sensors: [env * nsensordata] of the last substep's start state, or null
  float* sens_prev;          // accelerometers' previous frame velocities [env * (3 * nsens_acc + 1)], last: valid
  const int* order;          // env of each workgroup slot (a permutation of the batch; the identity by default)
};

// Sensors of the substep's start state (reference src/sensor.cpp:265-359),
// one lane per sensor, after the smooth forces: frame sensors from the body
// pose and cvel (point velocity about the root's subtree COM), joint and
// actuator sensors from qpos / qvel / actuator forces. Accelerometers
// differentiate their frame's world velocity over one substep (zero until a
// previous substep exists, gravity removed as in MuJoCo), so their velocity
// is kept every substep; the row is written on the last substep only.
KDEV void st_sensors(const DModel& m, const Layout& L, const float* w, const StepIO& io, int env, bool first,
                     bool last) {
  const float *xpos = w + L.xpos, *xquat = w + L.xquat, *xmat = w + L.xmat, *cvel = w + L.cvel, *scom = w + L.scom;
  const float *qpos = w + L.qpos, *qvel = w + L.qvel, *af = w + L.act_force;
  int pstride = 3 * m.nsens_acc + 1;
  float* prev = io.sens_prev + (size_t)env * pstride;
  // a previous velocity exists after the launch's first substep, and at the
  // first one when an earlier launch left it (a reset clears the flag)
  bool valid = m.nsens_acc > 0 && (!first || prev[pstride - 1] != 0.f);
  float* out = io.sensordata + (size_t)env * m.nsensordata;
  for (int i = LANE; i < m.nsensor; i += WAVE) {
    int t = m.sens_type[i], acc = m.sens_acc[i], b = m.sens_body[i];
    if (!last && acc < 0) continue;
    float y[4] = {0.f, 0.f, 0.f, 0.f};
    float pos[3] = {0.f, 0.f, 0.f}, R[9], q[4] = {1.f, 0.f, 0.f, 0.f}, ang[3] = {0.f, 0.f, 0.f}, lin[3] = {0.f, 0.f, 0.f};
    if (b >= 0 && t != SENS_SUBTREECOM) {  // the frame in world coordinates
      const float *Rb = xmat + 9 * b, *lp = m.sens_pos + 3 * i, *lq = m.sens_quat + 4 * i;
      float t3[3], Rl[9];
      mv3(Rb, lp, t3);
      for (int k = 0; k < 3; k++) pos[k] = xpos[3 * b + k] + t3[k];
      quat2mat(lq, Rl);
      for (int r = 0; r < 3; r++)
        for (int c = 0; c < 3; c++) R[3 * r + c] = Rb[3 * r] * Rl[c] + Rb[3 * r + 1] * Rl[3 + c] + Rb[3 * r + 2] * Rl[6 + c];
      mulquat(xquat + 4 * b, lq, q);
      const float *cv = cvel + 6 * b, *c = scom + 3 * m.body_rootid[b];
      float off[3] = {pos[0] - c[0], pos[1] - c[1], pos[2] - c[2]}, wx[3];
      for (int k = 0; k < 3; k++) ang[k] = cv[k];
      cross3(ang, off, wx);
      for (int k = 0; k < 3; k++) lin[k] = cv[3 + k] + wx[k];
    }
    auto rt = [&](const float* v, float* o) {  // R^T v
      for (int k = 0; k < 3; k++) o[k] = R[k] * v[0] + R[3 + k] * v[1] + R[6 + k] * v[2];
    };
    switch (t) {
      case SENS_JOINTPOS: y[0] = qpos[m.sens_qadr[i]]; break;
      case SENS_JOINTVEL: y[0] = qvel[m.sens_dadr[i]]; break;
      case SENS_ACTUATORPOS: y[0] = m.sens_gear[i] * qpos[m.sens_qadr[i]]; break;
      case SENS_ACTUATORVEL: y[0] = m.sens_gear[i] * qvel[m.sens_dadr[i]]; break;
      case SENS_JOINTACTFRC: {
        int k = m.sens_dadr[i];
        for (int a = m.dof_act_adr[k]; a < m.dof_act_adr[k] + m.dof_act_num[k]; a++) {
          int u = m.dof_act_list[a];
          y[0] += m.act_gear[u] * af[u];
        }
        break;
      }
      case SENS_BALLQUAT: {
        const float* qb = qpos + m.sens_qadr[i];
        float n = sqrtf(qb[0] * qb[0] + qb[1] * qb[1] + qb[2] * qb[2] + qb[3] * qb[3]);
        for (int k = 0; k < 4; k++) y[k] = qb[k] / n;
        break;
      }
      case SENS_BALLANGVEL: for (int k = 0; k < 3; k++) y[k] = qvel[m.sens_dadr[i] + k]; break;
      case SENS_SUBTREECOM: for (int k = 0; k < 3; k++) y[k] = scom[3 * b + k]; break;
      case SENS_FRAMEPOS: for (int k = 0; k < 3; k++) y[k] = pos[k]; break;
      case SENS_FRAMEQUAT: { float sg = q[0] < 0.f ? -1.f : 1.f; for (int k = 0; k < 4; k++) y[k] = sg * q[k]; break; }
      case SENS_FRAMEXAXIS: case SENS_FRAMEYAXIS: case SENS_FRAMEZAXIS: {
        int k = t - SENS_FRAMEXAXIS;
        y[0] = R[k]; y[1] = R[3 + k]; y[2] = R[6 + k];
        break;
      }
      case SENS_FRAMELINVEL: for (int k = 0; k < 3; k++) y[k] = lin[k]; break;
      case SENS_FRAMEANGVEL: for (int k = 0; k < 3; k++) y[k] = ang[k]; break;
      case SENS_GYRO: rt(ang, y); break;
      case SENS_VELOCIMETER: rt(lin, y); break;
      case SENS_MAGNETOMETER: rt(m.magnetic, y); break;
      case SENS_ACCELEROMETER: {
        float* pv = prev + 3 * acc;
        float a[3];
        for (int k = 0; k < 3; k++) a[k] = (valid ? (lin[k] - pv[k]) / m.timestep : 0.f) - m.gravity[k];
        rt(a, y);
        for (int k = 0; k < 3; k++) pv[k] = lin[k];
        break;
      }
    }
    if (last)
      for (int k = 0; k < m.sens_dim[i]; k++) out[m.sens_adr[i] + k] = y[k];
  }
  if (LANE == 0 && m.nsens_acc > 0 && last) prev[pstride - 1] = 1.f;
}

KDEV void dump_ws(const Layout& L, const float* w, float* dst, int env) {
  if (!dst) return;
  for (int t = LANE; t < L.total; t += WAVE) dst[(size_t)env * L.total + t] = w[t];
  KSYNC();
}

// Per-body contact force after the solve, [torque 3, force 3] per body: each
// contact's normal force (the sum of its rows' forces) along the contact
// normal, added to body2 and subtracted from body1; torque zero (reference
// src/compat/mkx_kernels/solver.hpp:906-929).
KDEV void write_cfrc_ext(const DModel& m, const Layout& L, const float* w, const AuxOut& aux, int env, bool solved) {
  int nb = m.nbody;
  const int* misc = reinterpret_cast<const int*>(w + L.misc);
  int ncon = solved ? misc[MISC_NCON] : 0, nefc = misc[MISC_NEFC];
  const float* force = w + L.efc_force;
  for (int b = LANE; b < nb; b += WAVE) {
    float f[3] = {0.f, 0.f, 0.f};
    for (int c = 0; c < ncon; c++) {
      const float* cr = w + L.con + c * CON_STRIDE;
      int p = reinterpret_cast<const int*>(cr + CON_PAIR)[0];
      float sgn = m.pair_b2[p] == b ? 1.f : (m.pair_b1[p] == b ? -1.f : 0.f);
      if (sgn == 0.f) continue;
      int nrow = reinterpret_cast<const int*>(cr + CON_NROW)[0], r0 = reinterpret_cast<const int*>(cr + CON_ROW0)[0];
      float fn = 0.f;
      if (m.cone != 0) fn = nrow > 0 && r0 < nefc ? force[r0] : 0.f;  // elliptic: the normal row
      else
        for (int r = r0; r < r0 + nrow && r < nefc; r++) fn += force[r];
      for (int k = 0; k < 3; k++) f[k] += sgn * fn * cr[CON_FRAME + k];
    }
    float* o = aux.cfrc_ext + ((size_t)env * nb + b) * 6;
    o[0] = o[1] = o[2] = 0.f;
    o[3] = f[0]; o[4] = f[1]; o[5] = f[2];
  }
}

KDEV void write_aux(const DModel& m, const Layout& L, const float* w, const AuxOut& aux, int env) {
  int nb = m.nbody;
  for (int t = LANE; t < 3 * nb; t += WAVE) {
    aux.xpos[(size_t)env * 3 * nb + t] = w[L.xpos + t];
    aux.subtree_com[(size_t)env * 3 * nb + t] = w[L.scom + t];
  }
  for (int t = LANE; t < 4 * nb; t += WAVE) aux.xquat[(size_t)env * 4 * nb + t] = w[L.xquat + t];
  for (int t = LANE; t < 10 * nb; t += WAVE) aux.cinert[(size_t)env * 10 * nb + t] = w[L.cinert + t];
  for (int t = LANE; t < 6 * nb; t += WAVE) aux.cvel[(size_t)env * 6 * nb + t] = w[L.cvel + t];
  for (int k = LANE; k < m.nv; k += WAVE) {
    float s = 0.f;
    for (int t = m.dof_act_adr[k]; t < m.dof_act_adr[k] + m.dof_act_num[k]; t++) {
      int u = m.dof_act_list[t];
      s += m.act_gear[u] * w[L.act_force + u];
    }
    aux.qfrc_actuator[(size_t)env * m.nv + k] = s;
  }
}

// nsub physics substeps of one env with the same ctrl; state stays in the
// workspace between substeps and goes back to HBM once at the end.
// Learned warm-start input of one env after a substep: qpos, qvel, ctrl
// (written as zeros here: the caller inserts the ctrl of the substep the
// prediction is for), then the first FEAT_NCON contacts as (normal, depth =
// -dist, one-hot of the limb class of either geom), zero-padded; with `prev`
// the substep's converged qacc follows (residual networks).
KDEV void write_features(const DModel& m, const Layout& L, const float* w, float* feat, int env, int prev) {
  int nq = m.nq, nv = m.nv, nu = m.nu, n0 = nq + nv + nu, nc = m.nfeat, nf = nc + (prev ? nv : 0);
  int ncon = reinterpret_cast<const int*>(w + L.misc)[MISC_NCON];
  for (int i = LANE; i < nf; i += WAVE) {
    float x = 0.f;
    if (i < nq) x = w[L.qpos + i];
    else if (i < nq + nv) x = w[L.qvel + i - nq];
    else if (i >= nc) x = w[L.qacc_ws + i - nc];
    else if (i >= n0) {
      int j = i - n0, c = j / FEAT_PER_CON, k = j - c * FEAT_PER_CON;
      if (c < ncon) {
        const float* cr = w + L.con + c * CON_STRIDE;
        if (k < 3) x = cr[CON_FRAME + k];
        else if (k == 3) x = -cr[CON_DIST];
        else {
          int p = reinterpret_cast<const int*>(cr + CON_PAIR)[0];
          int c1 = m.geom_fclass[m.pair_g1[p]], c2 = m.geom_fclass[m.pair_g2[p]];
          x = (c1 == k - 4 || c2 == k - 4) ? 1.f : 0.f;
        }
      }
    }
    feat[(size_t)env * nf + i] = x;
  }
}

// NPSET selects the narrow phases compiled in (0: plane / sphere / capsule
// pairs and box against plane and sphere; 1: also capsule-box and box-box;
// 2: also ellipsoids and cylinders through GJK / EPA),
// so models without the heavier pairs do not pay their registers.
// NPSET: bits 0-1 the narrow-phase set (0 basic pairs, 1 + capsule-box /
// box-box, 2 + GJK/EPA), bit 2 the sensor stage (models with sensors only, so
// the plain step carries none of its code), bit 3 set by the register-capped
// kernel variant
// m and L: the sim's device-memory copy (KParams), read through kfresh().
// lc: LDS for the line search's rows beside a device-memory workspace (or none)
template <int NPSET>
KDEV void run_env(const DModel& m, const Layout& L, float* w, const StepIO& io, int env, int nsub,
                  LsCache lc = LsCache{nullptr, 0}) {
  int nq = m.nq, nv = m.nv, nu = m.nu;
  for (int i = LANE; i < nq; i += WAVE) w[L.qpos + i] = io.qpos[(size_t)env * nq + i];
  for (int i = LANE; i < nv; i += WAVE) {
    w[L.qvel + i] = io.qvel[(size_t)env * nv + i];
    w[L.qacc_ws + i] = io.qacc_ws[(size_t)env * nv + i];
  }
  for (int i = LANE; i < nu; i += WAVE) w[L.ctrl + i] = io.ctrl ? io.ctrl[(size_t)env * nu + i] : 0.f;
  if (L.rand >= 0)
    for (int i = LANE; i < io.nrand; i += WAVE) w[L.rand + i] = io.rand[(size_t)env * io.nrand + i];
  KSYNC();
  if (nsub == 0) {
    // derived quantities of the current state only (no solve, no integration)
    st_kinematics(m, L, w);
    st_com_pos(m, L, w);
    st_crb(m, L, w);
    if (nv > 0) factor_ld<(NPSET & 8) == 0>(m, w + L.qM, w + L.qLD);
    st_com_vel(m, L, w);
    st_smooth(m, L, w);
    // the kinematics and dynamics intermediates are intact only here: the
    // constraint rows of a full substep take their space
    dump_ws(L, w, io.dump1, env);
    if (io.has_aux) {
      write_aux(m, L, w, io.aux, env);
      write_cfrc_ext(m, L, w, io.aux, env, false);
    }
    return;
  }
  int st = 0, niter = 0;
  KProfAcc pa;
  for (int i = 0; i < KP_N; i++) pa.t[i] = 0;
  KPROF_START(&pa);
  for (int s = 0; s < nsub; s++) {
    bool last = s == nsub - 1;
    st_kinematics(kfresh(m), kfresh(L), w);
    KPROF_MARK(&pa, KP_KIN);
    st_com_pos(kfresh(m), kfresh(L), w);
    KPROF_MARK(&pa, KP_COMPOS);
    st_crb(kfresh(m), kfresh(L), w);
    KPROF_MARK(&pa, KP_CRB);
    if (nv > 0) st |= factor_ld<(NPSET & 8) == 0>(kfresh(m), w + L.qM, w + L.qLD);
    KPROF_MARK(&pa, KP_FACTOR);
    st_com_vel(kfresh(m), kfresh(L), w);
    KPROF_MARK(&pa, KP_COMVEL);
    st |= st_smooth(kfresh(m), kfresh(L), w);
    KPROF_MARK(&pa, KP_SMOOTH);
    if (last && io.has_aux) write_aux(m, L, w, io.aux, env);
    if constexpr ((NPSET & 4) != 0)
      if (io.sensordata) st_sensors(kfresh(m), kfresh(L), w, io, env, s == 0, last);
    st |= st_collision<(NPSET & 3)>(kfresh(m), kfresh(L), w);
    KPROF_MARK(&pa, KP_COLL);
    st |= st_constraints(kfresh(m), kfresh(L), w);
    KPROF_MARK(&pa, KP_CONSTR);
    if (last) dump_ws(L, w, io.dump1, env);
    int it = 0;
    st |= st_solve<(NPSET & 8) == 0>(kfresh(m), kfresh(L), w, &it, &pa,
                                     io.qacc_pred && s == 0 ? io.qacc_pred + (size_t)env * nv : nullptr, lc);
    niter += it;
    if (last && io.has_aux) write_cfrc_ext(m, L, w, io.aux, env, true);
    if (last) dump_ws(L, w, io.dump2, env);
    st |= st_integrate<(NPSET & 8) == 0>(kfresh(m), kfresh(L), w);
    KPROF_MARK(&pa, KP_INTEGRATE);
  }
  int bad = 0;
  for (int i = LANE; i < nq; i += WAVE) {
    float v = w[L.qpos + i];
    if (!(v == v) || fabsf(v) > 3.0e38f) bad = 1;
    io.qpos[(size_t)env * nq + i] = v;
  }
  for (int i = LANE; i < nv; i += WAVE) {
    float v = w[L.qvel + i];
    if (!(v == v) || fabsf(v) > 3.0e38f) bad = 1;
    io.qvel[(size_t)env * nv + i] = v;
    io.qacc_ws[(size_t)env * nv + i] = w[L.qacc_ws + i];
  }
  if (wori(bad)) st |= ST_NONFINITE;
  if (io.feat) write_features(m, L, w, io.feat, env, io.feat_prev_qacc);
  KPROF_MARK(&pa, KP_IO);
  if (LANE == 0) {
    const int* misc = reinterpret_cast<const int*>(w + L.misc);
    if (io.prof)
      for (int i = 0; i < KP_N; i++) io.prof[(size_t)env * KP_N + i] = pa.t[i];
    if (io.niter) io.niter[env] = io.accumulate ? io.niter[env] + niter : niter;
    if (io.status) io.status[env] = io.accumulate ? (io.status[env] | st) : st;
    if (io.ncon) io.ncon[env] = misc[MISC_NCON_ALL];
    if (io.nefc) io.nefc[env] = misc[MISC_NEFC];
  }
}

}  // namespace mjk
