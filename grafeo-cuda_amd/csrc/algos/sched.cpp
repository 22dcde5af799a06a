// The learned scheduler object (an addition with no reference function): the
// FP32 MLP of mlp.hip resident on one device, its GPU inference, its loss
// gradient and its GPU training (full-batch Adam), and the hook that attaches
// it to a graph so that BFS picks each level's direction and triangle
// counting picks each degree bucket's intersection method with it. Without a
// scheduler the fixed rules run (the fallback).
#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "../runtime/common.hpp"

namespace gc {

void sched_release(Sched* s) {
  if (s == nullptr) return;
  if (s->refs.fetch_sub(1) == 1) {
    Device* dev = s->dev;
    {
      std::lock_guard<std::recursive_mutex> lock(dev->mu);
      (void)hipSetDevice(dev->ordinal);
      (void)hipStreamSynchronize(dev->stream);
      delete s;
    }
    gc_device_release(reinterpret_cast<gc_device*>(dev));
  }
}

void sched_predict(Sched& s, const float* x, uint32_t count, float* y) {
  if (count == 0) return;
  Device& dev = *s.dev;
  DeviceMemory buf;
  const size_t xb = sizeof(float) * count * kSchedFeatures, yb = sizeof(float) * count * kSchedChoices;
  buf.allocate(xb + yb);
  float* dx = buf.as<float>();
  float* dy = reinterpret_cast<float*>(buf.as<char>() + xb);
  copy_to_device(dev, dx, x, xb);
  const uint32_t blocks = std::max<uint32_t>(1, std::min<uint32_t>((count + 255) / 256, 4096));
  launch(dev, get_kernel(dev, "mlp", "mlp_forward"), blocks, 256u, 0, s.params.as<const float>(), s.norm,
         static_cast<const float*>(dx), count, dy);
  copy_to_host(dev, y, dy, yb);
}

namespace {

Sched* sched_from_handle(gc_sched* h) {
  if (h == nullptr) fail_kernel("scheduler is null");
  return reinterpret_cast<Sched*>(h);
}

struct SchedCall {
  Sched* s;
  std::unique_lock<std::recursive_mutex> lock;
  explicit SchedCall(gc_sched* h) : s(sched_from_handle(h)), lock(s->dev->mu) { s->dev->make_current(); }
};

uint64_t splitmix(uint64_t& state) {
  uint64_t z = (state += 0x9E3779B97F4A7C15ull);
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

// He-uniform weights (bound sqrt(6 / fan_in)), zero biases.
std::vector<float> initial_params(uint64_t seed) {
  std::vector<float> p(kSchedParams, 0.0f);
  uint64_t st = seed;
  auto uniform = [&](float bound) {
    const double u = static_cast<double>(splitmix(st) >> 11) * (1.0 / 9007199254740992.0);
    return static_cast<float>((2.0 * u - 1.0) * bound);
  };
  const uint32_t w1 = 0, b1 = w1 + 32 * 12, w2 = b1 + 32, b2 = w2 + 32 * 32, w3 = b2 + 32;
  for (uint32_t i = 0; i < 32 * 12; ++i) p[w1 + i] = uniform(std::sqrt(6.0f / 12.0f));
  for (uint32_t i = 0; i < 32 * 32; ++i) p[w2 + i] = uniform(std::sqrt(6.0f / 32.0f));
  for (uint32_t i = 0; i < 3 * 32; ++i) p[w3 + i] = uniform(std::sqrt(6.0f / 32.0f));
  return p;
}

struct TrainBuffers {
  float* x;
  float* t;
  float* mask;
  float* per_sample;
  float* losses;
  float inv_obs;
};

// Uploads a batch and sizes the per-sample gradient scratch.
TrainBuffers upload_batch(Sched& s, DeviceMemory& scratch, const float* x, const float* t,
                          const float* mask, uint32_t count) {
  Device& dev = *s.dev;
  double obs = 0.0;
  for (uint64_t i = 0; i < static_cast<uint64_t>(count) * kSchedChoices; ++i) obs += mask[i];
  if (obs <= 0.0) fail_kernel("training batch has no observed output (mask is all zero)");
  const size_t xb = sizeof(float) * count * kSchedFeatures, tb = sizeof(float) * count * kSchedChoices;
  const size_t gb = sizeof(float) * static_cast<size_t>(count) * kSchedParams, lb = sizeof(float) * count;
  scratch.ensure(xb + 2 * tb + gb + lb + 256);
  auto* base = scratch.as<char>();
  TrainBuffers b{reinterpret_cast<float*>(base), reinterpret_cast<float*>(base + xb),
                 reinterpret_cast<float*>(base + xb + tb), reinterpret_cast<float*>(base + xb + 2 * tb),
                 reinterpret_cast<float*>(base + xb + 2 * tb + gb), static_cast<float>(1.0 / obs)};
  copy_to_device(dev, b.x, x, xb);
  copy_to_device(dev, b.t, t, tb);
  copy_to_device(dev, b.mask, mask, tb);
  return b;
}

// Loss and full gradient of the uploaded batch into s.grad (loss after the
// kSchedParams gradient values).
void loss_grad(Sched& s, const TrainBuffers& b, uint32_t count) {
  Device& dev = *s.dev;
  const uint32_t blocks = std::max<uint32_t>(1, std::min<uint32_t>((count + 255) / 256, 4096));
  launch(dev, get_kernel(dev, "mlp", "mlp_backward"), blocks, 256u, 0,
         s.params.as<const float>(), s.norm, static_cast<const float*>(b.x),
         static_cast<const float*>(b.t), static_cast<const float*>(b.mask), count, b.inv_obs,
         b.per_sample, b.losses);
  launch(dev, get_kernel(dev, "mlp", "mlp_reduce"), (kSchedParams + 1 + 255) / 256, 256u, 0,
         static_cast<const float*>(b.per_sample), static_cast<const float*>(b.losses), count,
         s.grad.as<float>(), s.grad.as<float>() + kSchedParams);
}

}  // namespace

}  // namespace gc

using gc::guard;

extern "C" {

gc_status gc_sched_new(gc_device* dev_handle, uint64_t seed, gc_sched** out) {
  return guard([&] {
    if (out == nullptr) gc::fail_kernel("out is null");
    *out = nullptr;
    if (dev_handle == nullptr) gc::fail_kernel("device is null");
    auto* dev = reinterpret_cast<gc::Device*>(dev_handle);
    std::lock_guard<std::recursive_mutex> lock(dev->mu);
    dev->make_current();
    std::unique_ptr<gc::Sched> s(new gc::Sched());
    s->params.allocate(sizeof(float) * gc::kSchedParams);
    s->m1.allocate(sizeof(float) * gc::kSchedParams);
    s->m2.allocate(sizeof(float) * gc::kSchedParams);
    s->grad.allocate(sizeof(float) * (gc::kSchedParams + 1));
    const std::vector<float> p = gc::initial_params(seed);
    gc::copy_to_device(*dev, s->params.get(), p.data(), sizeof(float) * p.size());
    gc::fill_words(*dev, s->m1.get(), gc::kSchedParams, 0);
    gc::fill_words(*dev, s->m2.get(), gc::kSchedParams, 0);
    for (uint32_t i = 0; i < gc::kSchedFeatures; ++i) {
      s->norm.mean[i] = 0.0f;
      s->norm.scale[i] = 1.0f;
    }
    s->dev = reinterpret_cast<gc::Device*>(gc_device_retain(dev_handle));
    *out = reinterpret_cast<gc_sched*>(s.release());
  });
}

void gc_sched_free(gc_sched* h) {
  if (h != nullptr) gc::sched_release(reinterpret_cast<gc::Sched*>(h));
}

gc_status gc_sched_get(gc_sched* h, float* params, float* mean, float* scale) {
  return guard([&] {
    gc::SchedCall call(h);
    if (params) gc::copy_to_host(*call.s->dev, params, call.s->params.get(), sizeof(float) * gc::kSchedParams);
    if (mean) std::memcpy(mean, call.s->norm.mean, sizeof(call.s->norm.mean));
    if (scale) std::memcpy(scale, call.s->norm.scale, sizeof(call.s->norm.scale));
  });
}

gc_status gc_sched_set(gc_sched* h, const float* params, const float* mean, const float* scale) {
  return guard([&] {
    gc::SchedCall call(h);
    gc::Sched& s = *call.s;
    if (params) {
      for (uint32_t i = 0; i < gc::kSchedParams; ++i)
        if (!std::isfinite(params[i])) gc::fail_kernel("scheduler parameters must be finite");
      gc::copy_to_device(*s.dev, s.params.get(), params, sizeof(float) * gc::kSchedParams);
      gc::fill_words(*s.dev, s.m1.get(), gc::kSchedParams, 0);
      gc::fill_words(*s.dev, s.m2.get(), gc::kSchedParams, 0);
      s.adam_step = 0;
    }
    if (mean) std::memcpy(s.norm.mean, mean, sizeof(s.norm.mean));
    if (scale) std::memcpy(s.norm.scale, scale, sizeof(s.norm.scale));
    ++s.version;
    s.dev->sync();
  });
}

gc_status gc_sched_predict(gc_sched* h, const float* x, uint32_t count, float* y) {
  return guard([&] {
    gc::SchedCall call(h);
    if (count == 0) return;
    if (x == nullptr || y == nullptr) gc::fail_kernel("x or y is null");
    gc::sched_predict(*call.s, x, count, y);
  });
}

gc_status gc_sched_loss_grad(gc_sched* h, const float* x, const float* t, const float* mask,
                             uint32_t count, float* loss, float* grad) {
  return guard([&] {
    gc::SchedCall call(h);
    gc::Sched& s = *call.s;
    if (count == 0 || x == nullptr || t == nullptr || mask == nullptr) gc::fail_kernel("empty or null batch");
    gc::DeviceMemory scratch;
    const gc::TrainBuffers b = gc::upload_batch(s, scratch, x, t, mask, count);
    gc::loss_grad(s, b, count);
    std::vector<float> host(gc::kSchedParams + 1);
    gc::copy_to_host(*s.dev, host.data(), s.grad.get(), sizeof(float) * host.size());
    if (grad) std::memcpy(grad, host.data(), sizeof(float) * gc::kSchedParams);
    if (loss) *loss = host[gc::kSchedParams];
  });
}

gc_status gc_sched_train(gc_sched* h, const float* x, const float* t, const float* mask,
                         uint32_t count, uint32_t steps, float lr, float* final_loss) {
  return guard([&] {
    gc::SchedCall call(h);
    gc::Sched& s = *call.s;
    if (count == 0 || x == nullptr || t == nullptr || mask == nullptr) gc::fail_kernel("empty or null batch");
    if (!(lr > 0.0f)) gc::fail_kernel("learning rate must be positive");
    gc::Device& dev = *s.dev;
    gc::DeviceMemory scratch;
    const gc::TrainBuffers b = gc::upload_batch(s, scratch, x, t, mask, count);
    gc::Kernel adam = gc::get_kernel(dev, "mlp", "mlp_adam");
    for (uint32_t i = 0; i < steps; ++i) {
      gc::loss_grad(s, b, count);
      ++s.adam_step;
      ++s.version;
      gc::launch(dev, adam, (gc::kSchedParams + 255) / 256, 256u, 0, s.params.as<float>(),
                 static_cast<const float*>(s.grad.as<float>()), s.m1.as<float>(), s.m2.as<float>(),
                 s.adam_step, lr, 0.9f, 0.999f, 1e-8f);
    }
    gc::loss_grad(s, b, count);
    float l = 0.0f;
    gc::copy_to_host(dev, &l, s.grad.as<float>() + gc::kSchedParams, sizeof(float));
    if (final_loss) *final_loss = l;
  });
}

gc_status gc_graph_set_scheduler(gc_graph* gh, gc_sched* bfs, gc_sched* tc) {
  return guard([&] {
    gc::GraphCall call(gh);
    auto take = [&](gc_sched* h) -> gc::Sched* {
      if (h == nullptr) return nullptr;
      auto* s = reinterpret_cast<gc::Sched*>(h);
      if (s->dev != call.g->dev) gc::fail_kernel("scheduler and graph are on different devices");
      return s;
    };
    gc::Sched* b = take(bfs);
    gc::Sched* t = take(tc);
    if (b) b->refs.fetch_add(1);
    if (t) t->refs.fetch_add(1);
    gc::Sched* old_b = call.g->bfs_sched;
    gc::Sched* old_t = call.g->tc_sched;
    call.g->bfs_sched = b;
    call.g->tc_sched = t;
    call.g->tc_choice_sched = nullptr;  // decisions of another network are stale
    if (old_b != nullptr) gc::sched_release(old_b);
    if (old_t != nullptr) gc::sched_release(old_t);
  });
}

gc_status gc_graph_last_schedule(gc_graph* gh, uint32_t* choices, uint32_t capacity, uint32_t* count) {
  return guard([&] {
    gc::GraphCall call(gh);
    const std::vector<uint32_t>& sch = call.g->last_schedule;
    if (count) *count = static_cast<uint32_t>(sch.size());
    if (choices)
      for (uint32_t i = 0; i < capacity && i < sch.size(); ++i) choices[i] = sch[i];
  });
}

}  // extern "C"
