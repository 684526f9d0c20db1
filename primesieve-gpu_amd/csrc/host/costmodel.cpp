///
/// @file  costmodel.cpp
/// @brief Host side of the FP32 MLP cost model (kernels/psg_mlp.hip): device
///        buffers, minibatch training with the hand-written forward,
///        backward and Adam kernels, prediction, and the single-step entry
///        points the gradient checks call.  No reference counterpart (the
///        reference tunes by grid search only, tools/gpu_check.cpp:164-204).
///

#include "GpuSieve.hpp"
#include "HipBackend.hpp"

#include <primesieve/primesieve_error.hpp>
#include <primesieve_gpu.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <mutex>
#include <numeric>
#include <random>
#include <string>
#include <vector>

extern "C" const unsigned char psg_mlp_hsaco_start[];

namespace {

const uint32_t kParams = 4673;
const uint32_t kIn = 6;
const uint32_t kHidden = 64;
const uint32_t kTile = 32;

} // namespace

struct psg_mlp
{
  const psg::HipApi* api = nullptr;
  int device = 0;
  hipModule_t mod = nullptr;
  hipStream_t stream = nullptr;
  hipFunction_t fForward = nullptr, fBackward = nullptr, fAdam = nullptr, fZero = nullptr;
  psg::DeviceBuffer P, G, M, V, L;          // params, grads, Adam moments, loss
  psg::DeviceBuffer X, T, H1, H2, Y;        // data and activations
  size_t capacity = 0;                      // samples the data buffers hold
  uint32_t step = 0;                        // Adam steps taken

  ~psg_mlp()
  {
    if (!api) return;
    (void) api->SetDevice(device);
    P.release(); G.release(); M.release(); V.release(); L.release();
    X.release(); T.release(); H1.release(); H2.release(); Y.release();
    if (stream) (void) api->StreamDestroy(stream);
    if (mod) (void) api->ModuleUnload(mod);
  }

  void check(hipError_t e, const char* what) { psg::hipCheck(api, e, what); }

  void reserve(size_t n)
  {
    if (n <= capacity) return;
    X.allocate(api, n * kIn * 4);
    T.allocate(api, n * 4);
    H1.allocate(api, n * kHidden * 4);
    H2.allocate(api, n * kHidden * 4);
    Y.allocate(api, n * 4);
    capacity = n;
  }

  void launch(hipFunction_t f, uint32_t grid, uint32_t block, void** params, const char* name)
  {
    check(api->ModuleLaunchKernel(f, grid, 1, 1, block, 1, 1, 0, stream, params, nullptr), name);
  }

  void zero(psg::DeviceBuffer& b, uint32_t n)
  {
    float* p = (float*) b.get();
    void* a[] = { &p, &n };
    launch(fZero, (n + 255) / 256, 256, a, "psg_mlp_zero");
  }

  // forward over the samples at X + off (n of them), activations kept
  void forward(size_t off, uint32_t n)
  {
    const float* p = (const float*) P.get();
    const float* x = (const float*) X.get() + off * kIn;
    float* h1 = (float*) H1.get();
    float* h2 = (float*) H2.get();
    float* y = (float*) Y.get();
    void* a[] = { &p, &x, &n, &h1, &h2, &y };
    launch(fForward, (n + kTile - 1) / kTile, 64, a, "psg_mlp_forward");
  }

  // gradients of the samples at off (after forward on the same samples)
  void backward(size_t off, uint32_t n)
  {
    zero(G, kParams);
    const float* p = (const float*) P.get();
    const float* x = (const float*) X.get() + off * kIn;
    const float* t = (const float*) T.get() + off;
    const float* h1 = (const float*) H1.get();
    const float* h2 = (const float*) H2.get();
    const float* y = (const float*) Y.get();
    float* g = (float*) G.get();
    float* l = (float*) L.get();
    void* a[] = { &p, &x, &t, &h1, &h2, &y, &n, &g, &l };
    launch(fBackward, (n + kTile - 1) / kTile, 64, a, "psg_mlp_backward");
  }

  void adam(float lr)
  {
    step++;
    const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
    const float c1 = (float) (1.0 / (1.0 - std::pow(0.9, (double) step)));
    const float c2 = (float) (1.0 / (1.0 - std::pow(0.999, (double) step)));
    float* p = (float*) P.get();
    const float* g = (const float*) G.get();
    float* m = (float*) M.get();
    float* v = (float*) V.get();
    uint32_t n = kParams;
    void* a[] = { &p, &g, &m, &v, &n, (void*) &lr, (void*) &b1, (void*) &b2, (void*) &eps,
                  (void*) &c1, (void*) &c2 };
    launch(fAdam, (kParams + 255) / 256, 256, a, "psg_mlp_adam");
  }

  void sync() { check(api->StreamSynchronize(stream), "cost model kernels"); }
};

namespace {

void setError(const std::string& e)
{
  std::lock_guard<std::mutex> lock(primesieve::gpu::gpuMutex());
  primesieve::gpu::lastErrorRef() = e;
}

template <typename F>
int guarded(F f)
{
  try
  {
    f();
    return 0;
  }
  catch (const std::exception& e)
  {
    setError(e.what());
    return -1;
  }
}

} // namespace

extern "C" {

int psg_mlp_param_count(void) { return (int) kParams; }

psg_mlp* psg_mlp_create(void)
{
  psg_mlp* m = nullptr;
  int rc = guarded([&] {
    int device = 0;
    {
      std::lock_guard<std::mutex> lock(primesieve::gpu::gpuMutex());
      psg::HipBackend* b = primesieve::gpu::activeHipBackend();
      if (!b)
        throw primesieve::primesieve_error("GPU is not available: " + primesieve::gpu::lastErrorRef());
      device = b->device().deviceIndex;
    }
    std::string err;
    const psg::HipApi* api = psg::hipApi(err);
    std::unique_ptr<psg_mlp> o(new psg_mlp);
    o->api = api;
    o->device = device;
    o->check(api->SetDevice(device), "hipSetDevice");
    o->check(api->ModuleLoadData(&o->mod, psg_mlp_hsaco_start), "loading the cost-model code object");
    o->check(api->ModuleGetFunction(&o->fForward, o->mod, "psg_mlp_forward"), "psg_mlp_forward");
    o->check(api->ModuleGetFunction(&o->fBackward, o->mod, "psg_mlp_backward"), "psg_mlp_backward");
    o->check(api->ModuleGetFunction(&o->fAdam, o->mod, "psg_mlp_adam"), "psg_mlp_adam");
    o->check(api->ModuleGetFunction(&o->fZero, o->mod, "psg_mlp_zero"), "psg_mlp_zero");
    o->check(api->StreamCreateWithFlags(&o->stream, hipStreamNonBlocking), "hipStreamCreate");
    o->P.allocate(api, kParams * 4);
    o->G.allocate(api, kParams * 4);
    o->M.allocate(api, kParams * 4);
    o->V.allocate(api, kParams * 4);
    o->L.allocate(api, 4);
    o->zero(o->M, kParams);
    o->zero(o->V, kParams);
    o->zero(o->P, kParams);
    o->sync();
    m = o.release();
  });
  return rc == 0 ? m : nullptr;
}

void psg_mlp_destroy(psg_mlp* m) { delete m; }

int psg_mlp_set_params(psg_mlp* m, const float* params)
{
  if (!m || !params) return -1;
  return guarded([&] {
    m->check(m->api->SetDevice(m->device), "hipSetDevice");
    m->check(m->api->Memcpy(m->P.get(), params, kParams * 4, hipMemcpyHostToDevice), "hipMemcpy params");
    m->zero(m->M, kParams);
    m->zero(m->V, kParams);
    m->step = 0;
    m->sync();
  });
}

int psg_mlp_get_params(psg_mlp* m, float* params)
{
  if (!m || !params) return -1;
  return guarded([&] {
    m->check(m->api->SetDevice(m->device), "hipSetDevice");
    m->check(m->api->Memcpy(params, m->P.get(), kParams * 4, hipMemcpyDeviceToHost), "hipMemcpy params");
  });
}

int psg_mlp_forward(psg_mlp* m, const float* x, size_t n, float* y)
{
  if (!m || (n && (!x || !y))) return -1;
  return guarded([&] {
    if (!n) return;
    m->check(m->api->SetDevice(m->device), "hipSetDevice");
    m->reserve(n);
    m->check(m->api->Memcpy(m->X.get(), x, n * kIn * 4, hipMemcpyHostToDevice), "hipMemcpy x");
    m->forward(0, (uint32_t) n);
    m->sync();
    m->check(m->api->Memcpy(y, m->Y.get(), n * 4, hipMemcpyDeviceToHost), "hipMemcpy y");
  });
}

int psg_mlp_gradients(psg_mlp* m, const float* x, const float* t, size_t n, float* grads, float* loss)
{
  if (!m || !n || !x || !t || !grads || !loss) return -1;
  return guarded([&] {
    m->check(m->api->SetDevice(m->device), "hipSetDevice");
    m->reserve(n);
    m->check(m->api->Memcpy(m->X.get(), x, n * kIn * 4, hipMemcpyHostToDevice), "hipMemcpy x");
    m->check(m->api->Memcpy(m->T.get(), t, n * 4, hipMemcpyHostToDevice), "hipMemcpy t");
    m->zero(m->L, 1);
    m->forward(0, (uint32_t) n);
    m->backward(0, (uint32_t) n);
    m->sync();
    m->check(m->api->Memcpy(grads, m->G.get(), kParams * 4, hipMemcpyDeviceToHost), "hipMemcpy grads");
    m->check(m->api->Memcpy(loss, m->L.get(), 4, hipMemcpyDeviceToHost), "hipMemcpy loss");
  });
}

int psg_mlp_adam_step(psg_mlp* m, const float* grads, float lr)
{
  if (!m || !grads) return -1;
  return guarded([&] {
    m->check(m->api->SetDevice(m->device), "hipSetDevice");
    m->check(m->api->Memcpy(m->G.get(), grads, kParams * 4, hipMemcpyHostToDevice), "hipMemcpy grads");
    m->adam(lr);
    m->sync();
  });
}

int psg_mlp_train(psg_mlp* m, const float* x, const float* t, size_t n, uint32_t epochs,
                  uint32_t batch, float lr, uint32_t seed, float* epoch_loss)
{
  if (!m || !n || !x || !t || !batch) return -1;
  return guarded([&] {
    m->check(m->api->SetDevice(m->device), "hipSetDevice");
    m->reserve(n);
    std::vector<uint32_t> perm(n);
    std::iota(perm.begin(), perm.end(), 0u);
    std::mt19937_64 rng(seed);
    std::vector<float> xs(n * kIn), ts(n);
    for (uint32_t e = 0; e < epochs; e++)
    {
      std::shuffle(perm.begin(), perm.end(), rng);
      for (size_t i = 0; i < n; i++)
      {
        std::memcpy(&xs[i * kIn], &x[(size_t) perm[i] * kIn], kIn * 4);
        ts[i] = t[perm[i]];
      }
      m->check(m->api->MemcpyAsync(m->X.get(), xs.data(), n * kIn * 4, hipMemcpyHostToDevice, m->stream),
               "hipMemcpyAsync x");
      m->check(m->api->MemcpyAsync(m->T.get(), ts.data(), n * 4, hipMemcpyHostToDevice, m->stream),
               "hipMemcpyAsync t");
      m->zero(m->L, 1);
      uint32_t batches = 0;
      for (size_t off = 0; off < n; off += batch, batches++)
      {
        const uint32_t b = (uint32_t) std::min<size_t>(batch, n - off);
        m->forward(off, b);
        m->backward(off, b);
        m->adam(lr);
      }
      m->sync();   // the host buffers are refilled next epoch
      if (epoch_loss)
      {
        float l = 0;
        m->check(m->api->Memcpy(&l, m->L.get(), 4, hipMemcpyDeviceToHost), "hipMemcpy loss");
        epoch_loss[e] = l / (float) batches;
      }
    }
  });
}

} // extern "C"
