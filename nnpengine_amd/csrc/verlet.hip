// verlet.hip -- velocity Verlet (VERLET, src/verlet.cpp:29-97).
//
// FP32 form: the reference's update, operation for operation:
//   updateX: x += (v + a*hdt)*dt; a_prev = a        updateV: a = f*(1/m)*accelF; v += (a_prev + a)*hdt
// (the force array holds the acceleration after updateV, as in the reference).
// FP64 form: x, v, a in double; the FP32 force is widened before the update and the FP32
// positions the force kernels read are re-derived from the FP64 master copy.
#include "internal.h"

namespace {

constexpr float kAccelF = 4.184e-4f;          // src/environ.h:78
constexpr double kEkinF = 1195.0286802753603; // src/environ.h:79

// the MD step's skin check on the updated position (xprev non-null): a wave with a row that
// moved more than skin/2 since the last build sets the invalid flag
__device__ __forceinline__ void skin_flag(bool bad, DevCounters* cnt) {
  if (__any(bad) && (threadIdx.x & 63) == 0) atomicOr(&cnt->invalid, 1);
}

__device__ __forceinline__ bool moved_too_far(float4 a, float4 b, float dx2) {
  const float dx = a.x - b.x, dy = a.y - b.y, dz = a.z - b.z;
  return dx * dx + dy * dy + dz * dz > dx2;
}

// skip (optional): rows of boxes flagged there keep their positions (a box's first step)
__global__ void k_update_x32(int n, float dt, float hdt, float4* xq, const float4* v, const float4* f, float4* fprev,
                             const int* rbox, const unsigned char* skip, const float4* xprev, float dx2, DevCounters* cnt) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  bool bad = false;
  if (i < n && !(skip && skip[rbox[i]])) {
    float4 x = xq[i];
    const float4 vv = v[i], ff = f[i];
    x.x += (vv.x + ff.x * hdt) * dt;
    x.y += (vv.y + ff.y * hdt) * dt;
    x.z += (vv.z + ff.z * hdt) * dt;
    xq[i] = x;
    fprev[i] = ff;
    if (xprev) bad = moved_too_far(x, xprev[i], dx2);
  }
  if (xprev) skin_flag(bad, cnt);
}

__global__ void k_update_v32(int n, float hdt, const float* rm, float4* v, float4* f, const float4* fprev) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float r = rm[i] * kAccelF;
  float4 ff = f[i], vv = v[i];
  const float4 fp = fprev[i];
  ff.x *= r; ff.y *= r; ff.z *= r;
  vv.x += (fp.x + ff.x) * hdt;
  vv.y += (fp.y + ff.y) * hdt;
  vv.z += (fp.z + ff.z) * hdt;
  f[i] = ff;
  v[i] = vv;
}

__global__ void k_update_x64(int n, double dt, double4* xd, const double4* vd, double4* ad, const float4* f, float4* xq,
                             const int* rbox, const unsigned char* skip, const float4* xprev, float dx2, DevCounters* cnt) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  bool bad = false;
  if (i < n && !(skip && skip[rbox[i]])) {
    double4 x = xd[i];
    const double4 vv = vd[i], a = ad[i];
    const double hdt = 0.5 * dt;
    x.x += (vv.x + a.x * hdt) * dt;
    x.y += (vv.y + a.y * hdt) * dt;
    x.z += (vv.z + a.z * hdt) * dt;
    xd[i] = x;
    float4 q = xq[i];
    q.x = float(x.x); q.y = float(x.y); q.z = float(x.z);
    xq[i] = q;
    if (xprev) bad = moved_too_far(q, xprev[i], dx2);
  }
  if (xprev) skin_flag(bad, cnt);
  (void)f;
}

// a_new = f/m*accelF (double); v += (a_prev + a_new)*hdt; a_prev <- a_new
__global__ void k_update_v64(int n, double dt, const float* rm, double4* vd, double4* ad, float4* f) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  f[i] = vv_kick64(i, dt, rm, vd, ad, f[i]);
}

__global__ void k_ke(int n, bool fp64, const float* rm, const float4* v, const double4* vd, double* part, double* part2) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  double ke = 0;
  if (i < n) {
    if (fp64) { double4 w = vd[i]; ke = (w.x * w.x + w.y * w.y + w.z * w.z) / double(rm[i]); }
    else { float4 w = v[i]; ke = double((w.x * w.x + w.y * w.y + w.z * w.z) / rm[i]); }
  }
  block_sum2_store<256>(ke * kEkinF, 0.0, part, part2);
}

__global__ void k_ke_final(const double* part, int nb, double* out) {
  double a = 0;
  for (int k = threadIdx.x; k < nb; k += 64) a += part[k];
  for (int o = 32; o > 0; o >>= 1) a += __shfl_down(a, o, 64);
  if (threadIdx.x == 0) *out = a;
}

__global__ void k_scale_v(int n, bool fp64, double s, float4* v, double4* vd) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  if (fp64) { double4 w = vd[i]; w.x *= s; w.y *= s; w.z *= s; vd[i] = w; }
  else { float4 w = v[i]; float fs = float(s); w.x *= fs; w.y *= fs; w.z *= fs; v[i] = w; }
}

__global__ void k_clear_f(int n, float4* f) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) f[i] = make_float4(0, 0, 0, 0);
}

// the integration state changes representation: copy it across so a run continues seamlessly
__global__ void k_to_fp64(int n, const float4* xq, const float4* v, const float4* fprev, double4* xd, double4* vd,
                          double4* ad) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float4 x = xq[i], w = v[i], a = fprev[i];
  xd[i] = make_double4(x.x, x.y, x.z, 0);
  vd[i] = make_double4(w.x, w.y, w.z, 0);
  ad[i] = make_double4(a.x, a.y, a.z, 0);
}

__global__ void k_to_fp32(int n, float4* xq, float4* v, float4* fprev, const double4* xd, const double4* vd,
                          const double4* ad) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double4 x = xd[i], w = vd[i], a = ad[i];
  float4 q = xq[i];
  q.x = float(x.x); q.y = float(x.y); q.z = float(x.z);
  xq[i] = q;
  v[i] = make_float4(float(w.x), float(w.y), float(w.z), 0);
  fprev[i] = make_float4(float(a.x), float(a.y), float(a.z), 0);
}

}  // namespace

int verlet_switch_precision(nnpe_ctx* ctx, bool fp64) {
  if (fp64 == ctx->fp64) return NNPE_OK;
  const int N = ctx->natom;
  if (fp64)
    hipLaunchKernelGGL(k_to_fp64, dim3(blocks_for(N, 256)), dim3(256), 0, ctx->stream, N, ctx->xq, ctx->v, ctx->fprev, ctx->xd,
                       ctx->vd, ctx->ad);
  else
    hipLaunchKernelGGL(k_to_fp32, dim3(blocks_for(N, 256)), dim3(256), 0, ctx->stream, N, ctx->xq, ctx->v, ctx->fprev, ctx->xd,
                       ctx->vd, ctx->ad);
  NNPE_CHECK_LAUNCH(ctx);
  ctx->fp64 = fp64;
  return NNPE_OK;
}

extern "C" int nnpe_verlet_set(nnpe_ctx* ctx, double dt, int fp64) {
  if (!ctx) return NNPE_ERR_ARG;
  if (!ctx->have_system) return ctx_fail(ctx, NNPE_ERR_STATE, "no system");
  md_state_free(ctx);
  ctx->dt = dt;
  return verlet_switch_precision(ctx, fp64 != 0);
}

static int update_x_enqueue(nnpe_ctx* ctx, const unsigned char* box_skip, bool skin) {
  const int N = ctx->natom;
  const float4* xprev = skin ? ctx->xprev : nullptr;
  if (ctx->fp64)
    hipLaunchKernelGGL(k_update_x64, dim3(blocks_for(N, 256)), dim3(256), 0, ctx->stream, N, ctx->dt, ctx->xd, ctx->vd, ctx->ad,
                       ctx->f, ctx->xq, ctx->rbox, box_skip, xprev, ctx->dx2, ctx->counters);
  else {
    const float dt = float(ctx->dt), hdt = dt * .5f;  // VERLET::setDt (src/verlet.cpp:206-210)
    hipLaunchKernelGGL(k_update_x32, dim3(blocks_for(N, 256)), dim3(256), 0, ctx->stream, N, dt, hdt, ctx->xq, ctx->v, ctx->f,
                       ctx->fprev, ctx->rbox, box_skip, xprev, ctx->dx2, ctx->counters);
  }
  NNPE_CHECK_LAUNCH(ctx);
  return NNPE_OK;
}

int verlet_enqueue_x_skip(nnpe_ctx* ctx, const unsigned char* box_skip) { return update_x_enqueue(ctx, box_skip, false); }

int verlet_enqueue_x(nnpe_ctx* ctx) { return update_x_enqueue(ctx, nullptr, false); }

int verlet_enqueue_x_skin(nnpe_ctx* ctx) { return update_x_enqueue(ctx, nullptr, true); }

int verlet_enqueue_v(nnpe_ctx* ctx) {
  const int N = ctx->natom;
  if (ctx->fp64)
    hipLaunchKernelGGL(k_update_v64, dim3(blocks_for(N, 256)), dim3(256), 0, ctx->stream, N, ctx->dt, ctx->rm, ctx->vd, ctx->ad,
                       ctx->f);
  else {
    const float hdt = float(ctx->dt) * .5f;
    hipLaunchKernelGGL(k_update_v32, dim3(blocks_for(N, 256)), dim3(256), 0, ctx->stream, N, hdt, ctx->rm, ctx->v, ctx->f,
                       ctx->fprev);
  }
  NNPE_CHECK_LAUNCH(ctx);
  return NNPE_OK;
}

int ke_enqueue(nnpe_ctx* ctx) {
  const int N = ctx->natom, nb = blocks_for(N, 256);
  double* part = ev_part(ctx, NNPE_TERM_KE, 0);
  hipLaunchKernelGGL(k_ke, dim3(nb), dim3(256), 0, ctx->stream, N, ctx->fp64, ctx->rm, ctx->v, ctx->vd, part,
                     ev_part(ctx, NNPE_TERM_KE, 1));
  hipLaunchKernelGGL(k_ke_final, dim3(1), dim3(64), 0, ctx->stream, part, nb, ctx->ev_sum + 2 * NNPE_TERM_KE);
  NNPE_CHECK_LAUNCH(ctx);
  return NNPE_OK;
}

extern "C" int nnpe_verlet_update_x(nnpe_ctx* ctx) {
  if (!ctx) return NNPE_ERR_ARG;
  if (!ctx->have_system) return ctx_fail(ctx, NNPE_ERR_STATE, "no system");
  return verlet_enqueue_x(ctx);
}

extern "C" int nnpe_verlet_update_v(nnpe_ctx* ctx) {
  if (!ctx) return NNPE_ERR_ARG;
  if (!ctx->have_system) return ctx_fail(ctx, NNPE_ERR_STATE, "no system");
  return verlet_enqueue_v(ctx);
}

extern "C" int nnpe_verlet_compute_ke(nnpe_ctx* ctx, double* ke) {
  if (!ctx || !ke) return NNPE_ERR_ARG;
  int rc = ke_enqueue(ctx);
  if (rc) return rc;
  double out[2 * NNPE_NTERM];
  rc = nnpe_energy_get(ctx, out);
  if (rc) return rc;
  *ke = out[2 * NNPE_TERM_KE];
  return NNPE_OK;
}

extern "C" int nnpe_verlet_scale_v(nnpe_ctx* ctx, double s) {
  if (!ctx) return NNPE_ERR_ARG;
  hipLaunchKernelGGL(k_scale_v, dim3(blocks_for(ctx->natom, 256)), dim3(256), 0, ctx->stream, ctx->natom, ctx->fp64, s, ctx->v,
                     ctx->vd);
  NNPE_CHECK_LAUNCH(ctx);
  return NNPE_OK;
}

extern "C" int nnpe_clear_force(nnpe_ctx* ctx) {
  if (!ctx) return NNPE_ERR_ARG;
  hipLaunchKernelGGL(k_clear_f, dim3(blocks_for(ctx->natom, 256)), dim3(256), 0, ctx->stream, ctx->natom, ctx->f);
  NNPE_CHECK_LAUNCH(ctx);
  return NNPE_OK;
}
