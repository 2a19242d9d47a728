// Host side of the task cost model: parameter init, batch upload, the
// captured training step and inference (kernels in xt_mlp.hip).
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/exachem_t.h"
#include "xt_device.h"
#include "xt_internal.h"

#define XT_HIP(call)                                                                    \
  do {                                                                                  \
    hipError_t e_ = (call);                                                             \
    if(e_ != hipSuccess)                                                                \
      return xt::set_error(XT_EHIP, std::string(#call) + ": " + hipGetErrorString(e_));  \
  } while(0)

static_assert(xt::kMlpParams == XT_MLP_PARAMS && xt::kMlpIn == XT_MLP_IN &&
                xt::kMlpHidden == XT_MLP_HIDDEN,
              "public and device model sizes agree");

struct xt_mlp {
  int         device = 0;
  hipStream_t stream = nullptr;
  xt::MlpWork W      = {};
  long        cap    = 0;  // batch capacity of X/T/Y/A*/D*/DY/SQ
  float*      hist   = nullptr;
  int         hist_cap = 0;

  void free_batch() {
    float** ps[] = {&W.X, &W.T, &W.Y, &W.A1, &W.A2, &W.D1, &W.D2, &W.DY, &W.SQ};
    for(float** p: ps)
      if(*p) {
        (void) hipFree(*p);
        *p = nullptr;
      }
    cap = 0;
  }
};

namespace {

inline uint64_t mix(uint64_t x) {
  x += 0x9E3779B97F4A7C15ULL;
  x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ULL;
  x = (x ^ (x >> 27)) * 0x94D049BB133111EBULL;
  return x ^ (x >> 31);
}
inline double unit(uint64_t h) { return ((double) (h >> 11) + 0.5) * (1.0 / 9007199254740992.0); }
inline double normal(uint64_t seed, uint64_t i) {
  const double u1 = unit(mix(seed ^ mix(2 * i))), u2 = unit(mix(seed ^ mix(2 * i + 1)));
  return std::sqrt(-2.0 * std::log(u1)) * std::cos(6.283185307179586 * u2);
}

int ensure_batch(xt_mlp* m, long n) {
  if(n <= m->cap) return XT_OK;
  m->free_batch();
  const long H = xt::kMlpHidden;
  auto al = [](float** p, long k) { return hipMalloc((void**) p, sizeof(float) * (size_t) k); };
  XT_HIP(al(&m->W.X, n * xt::kMlpIn));
  XT_HIP(al(&m->W.T, n));
  XT_HIP(al(&m->W.Y, n));
  XT_HIP(al(&m->W.A1, n * H));
  XT_HIP(al(&m->W.A2, n * H));
  XT_HIP(al(&m->W.D1, n * H));
  XT_HIP(al(&m->W.D2, n * H));
  XT_HIP(al(&m->W.DY, n));
  XT_HIP(al(&m->W.SQ, n));
  m->cap = n;
  return XT_OK;
}

int upload_batch(xt_mlp* m, long n, const float* x, const float* t) {
  int rc = ensure_batch(m, n);
  if(rc) return rc;
  XT_HIP(hipMemcpyAsync(m->W.X, x, sizeof(float) * n * xt::kMlpIn, hipMemcpyHostToDevice, m->stream));
  if(t) XT_HIP(hipMemcpyAsync(m->W.T, t, sizeof(float) * n, hipMemcpyHostToDevice, m->stream));
  return XT_OK;
}

int reset_adam(xt_mlp* m) {
  XT_HIP(hipMemsetAsync(m->W.M1, 0, sizeof(float) * xt::kMlpParams, m->stream));
  XT_HIP(hipMemsetAsync(m->W.M2, 0, sizeof(float) * xt::kMlpParams, m->stream));
  XT_HIP(hipMemsetAsync(m->W.step, 0, sizeof(int) * 2, m->stream));
  return XT_OK;
}

}  // namespace

extern "C" {

int xt_mlp_create(int device, uint64_t seed, xt_mlp** out) {
  if(!out) return xt::set_error(XT_EINVAL, "xt_mlp_create: out is null");
  *out  = nullptr;
  int n = 0;
  if(hipGetDeviceCount(&n) != hipSuccess || n <= 0)
    return xt::set_error(XT_ENODEV, "xt_mlp_create: no HIP device");
  if(device < 0 || device >= n) return xt::set_error(XT_EINVAL, "xt_mlp_create: bad device");
  auto* m   = new xt_mlp();
  m->device = device;
  XT_HIP(hipSetDevice(device));
  XT_HIP(hipStreamCreateWithFlags(&m->stream, hipStreamNonBlocking));
  const size_t ps = sizeof(float) * xt::kMlpParams;
  XT_HIP(hipMalloc((void**) &m->W.P, ps));
  XT_HIP(hipMalloc((void**) &m->W.G, ps));
  XT_HIP(hipMalloc((void**) &m->W.M1, ps));
  XT_HIP(hipMalloc((void**) &m->W.M2, ps));
  XT_HIP(hipMalloc((void**) &m->W.loss, sizeof(float)));
  XT_HIP(hipMalloc((void**) &m->W.step, 2 * sizeof(int)));
  // He initialisation: W1 ~ N(0, 2/7), W2 ~ N(0, 2/64), w3 ~ N(0, 1/64), biases 0
  const int           IN = xt::kMlpIn, H = xt::kMlpHidden;
  std::vector<float>  p(xt::kMlpParams, 0.f);
  uint64_t            k = 0;
  for(int i = 0; i < IN * H; i++) p[i] = (float) (std::sqrt(2.0 / IN) * normal(seed, k++));
  const int ow2 = IN * H + H;
  for(int i = 0; i < H * H; i++) p[ow2 + i] = (float) (std::sqrt(2.0 / H) * normal(seed, k++));
  const int ow3 = ow2 + H * H + H;
  for(int i = 0; i < H; i++) p[ow3 + i] = (float) (std::sqrt(1.0 / H) * normal(seed, k++));
  *out   = m;
  int rc = xt_mlp_set_params(m, p.data());
  if(rc) {
    xt_mlp_destroy(m);
    *out = nullptr;
  }
  return rc;
}

int xt_mlp_destroy(xt_mlp* m) {
  if(!m) return XT_OK;
  (void) hipSetDevice(m->device);
  if(m->stream) (void) hipStreamSynchronize(m->stream);
  m->free_batch();
  void* ps[] = {m->W.P, m->W.G, m->W.M1, m->W.M2, m->W.loss, m->W.step, m->hist};
  for(void* p: ps)
    if(p) (void) hipFree(p);
  if(m->stream) (void) hipStreamDestroy(m->stream);
  delete m;
  return XT_OK;
}

int xt_mlp_get_params(xt_mlp* m, float* params) {
  if(!m || !params) return xt::set_error(XT_EINVAL, "xt_mlp_get_params: null argument");
  XT_HIP(hipSetDevice(m->device));
  XT_HIP(hipMemcpyAsync(params, m->W.P, sizeof(float) * xt::kMlpParams, hipMemcpyDeviceToHost,
                        m->stream));
  XT_HIP(hipStreamSynchronize(m->stream));
  return XT_OK;
}

int xt_mlp_set_params(xt_mlp* m, const float* params) {
  if(!m || !params) return xt::set_error(XT_EINVAL, "xt_mlp_set_params: null argument");
  XT_HIP(hipSetDevice(m->device));
  XT_HIP(hipMemcpyAsync(m->W.P, params, sizeof(float) * xt::kMlpParams, hipMemcpyHostToDevice,
                        m->stream));
  int rc = reset_adam(m);
  if(rc) return rc;
  XT_HIP(hipStreamSynchronize(m->stream));
  return XT_OK;
}

int xt_mlp_loss_grad(xt_mlp* m, long n, const float* x, const float* t, float* loss, float* grad) {
  if(!m || n <= 0 || !x || !t) return xt::set_error(XT_EINVAL, "xt_mlp_loss_grad: bad argument");
  XT_HIP(hipSetDevice(m->device));
  int rc = upload_batch(m, n, x, t);
  if(rc) return rc;
  int err = xt::mlp_loss_grad(m->W, n, m->stream);
  if(err) return xt::set_error(XT_EHIP, std::string("mlp kernels: ") + hipGetErrorString((hipError_t) err));
  if(loss) XT_HIP(hipMemcpyAsync(loss, m->W.loss, sizeof(float), hipMemcpyDeviceToHost, m->stream));
  if(grad)
    XT_HIP(hipMemcpyAsync(grad, m->W.G, sizeof(float) * xt::kMlpParams, hipMemcpyDeviceToHost,
                          m->stream));
  XT_HIP(hipStreamSynchronize(m->stream));
  return XT_OK;
}

int xt_mlp_train(xt_mlp* m, long n, const float* x, const float* t, int steps, float lr,
                 float* loss_hist) {
  if(!m || n <= 0 || !x || !t || steps < 0 || !(lr > 0.f))
    return xt::set_error(XT_EINVAL, "xt_mlp_train: bad argument");
  if(steps == 0) return XT_OK;
  XT_HIP(hipSetDevice(m->device));
  int rc = upload_batch(m, n, x, t);
  if(rc) return rc;
  if(loss_hist && m->hist_cap < steps) {
    if(m->hist) (void) hipFree(m->hist);
    m->hist     = nullptr;
    m->hist_cap = 0;
    XT_HIP(hipMalloc((void**) &m->hist, sizeof(float) * steps));
    m->hist_cap = steps;
  }
  float* hist = loss_hist ? m->hist : nullptr;
  // step[1] = step[0]: base of this call's loss history
  XT_HIP(hipMemcpyAsync(m->W.step + 1, m->W.step, sizeof(int), hipMemcpyDeviceToDevice, m->stream));
  // one Adam step captured once, replayed `steps` times
  XT_HIP(hipStreamBeginCapture(m->stream, hipStreamCaptureModeThreadLocal));
  int        err = xt::mlp_train_step(m->W, n, lr, hist, m->stream);
  hipGraph_t g   = nullptr;
  hipError_t e   = hipStreamEndCapture(m->stream, &g);
  if(err || e != hipSuccess) {
    if(g) (void) hipGraphDestroy(g);
    return xt::set_error(XT_EHIP, std::string("mlp training capture: ") +
                                    hipGetErrorString(err ? (hipError_t) err : e));
  }
  hipGraphExec_t ge = nullptr;
  e                 = hipGraphInstantiate(&ge, g, nullptr, nullptr, 0);
  (void) hipGraphDestroy(g);
  if(e != hipSuccess) return xt::set_error(XT_EHIP, std::string("mlp graph: ") + hipGetErrorString(e));
  for(int s = 0; s < steps && e == hipSuccess; s++) e = hipGraphLaunch(ge, m->stream);
  hipError_t e2 = hipStreamSynchronize(m->stream);
  (void) hipGraphExecDestroy(ge);
  if(e != hipSuccess || e2 != hipSuccess)
    return xt::set_error(XT_EHIP, std::string("mlp training: ") +
                                    hipGetErrorString(e != hipSuccess ? e : e2));
  if(loss_hist) XT_HIP(hipMemcpy(loss_hist, m->hist, sizeof(float) * steps, hipMemcpyDeviceToHost));
  return XT_OK;
}

int xt_mlp_predict(xt_mlp* m, long n, const float* x, float* y) {
  if(!m || n < 0 || (n > 0 && (!x || !y))) return xt::set_error(XT_EINVAL, "xt_mlp_predict: bad argument");
  if(n == 0) return XT_OK;
  XT_HIP(hipSetDevice(m->device));
  int rc = upload_batch(m, n, x, nullptr);
  if(rc) return rc;
  int err = xt::mlp_forward(m->W.P, m->W.X, n, m->W.Y, nullptr, nullptr, m->stream);
  if(err) return xt::set_error(XT_EHIP, std::string("mlp forward: ") + hipGetErrorString((hipError_t) err));
  XT_HIP(hipMemcpyAsync(y, m->W.Y, sizeof(float) * n, hipMemcpyDeviceToHost, m->stream));
  XT_HIP(hipStreamSynchronize(m->stream));
  return XT_OK;
}

}  // extern "C"
