// Host stand-ins for the HIP runtime calls and the bake entry points that the bake server (ib_server.cpp)
// uses, so that the server's own logic -- admission by budget, the request queues, the worker threads, the
// resident-mesh cache and its eviction, fetch and release -- runs under the host sanitizers without a GPU.
// "Device" memory is host memory.  A bake writes a deterministic function of its soup (fake_expected_*),
// which the driver recomputes; a soup of exactly 13 triangles makes the bake fail.
#include <hip/hip_runtime_api.h>

#include <chrono>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "fake_device.h"
#include "italy_bake.h"

hipError_t hipMalloc(void **p, size_t n) {
  *p = std::malloc(n ? n : 1);
  return *p ? hipSuccess : hipErrorOutOfMemory;
}
hipError_t hipFree(void *p) {
  std::free(p);
  return hipSuccess;
}
hipError_t hipMemcpy(void *dst, const void *src, size_t n, hipMemcpyKind) {
  std::memcpy(dst, src, n);
  return hipSuccess;
}
hipError_t hipSetDevice(int d) { return d == 0 ? hipSuccess : hipErrorInvalidDevice; }
hipError_t hipMemGetInfo(size_t *freeb, size_t *total) {
  *freeb = 1ull << 30;
  *total = 1ull << 31;
  return hipSuccess;
}

struct ib_context {
  std::mutex mu;  // one bake at a time per context, as in the library
};
struct ib_mesh {
  ib_context *ctx;
  std::vector<float> soup;
};

namespace {
thread_local std::string g_err;
}

uint64_t fake_soup_key(const float *xyz, uint64_t ntri) {
  uint64_t h = 1469598103934665603ull;
  for (uint64_t i = 0; i < 9 * ntri; ++i) {
    uint32_t b;
    std::memcpy(&b, &xyz[i], 4);
    h = (h ^ b) * 1099511628211ull;
  }
  return h;
}
float fake_expected_distance(uint64_t key, uint64_t i) { return (float)((key + i) % 100003); }
bool fake_expected_brick(uint64_t key, uint64_t b) { return (key + b) % 3 == 0; }
uint32_t fake_expected_word(uint64_t key, uint64_t i) { return (uint32_t)(key * 2654435761ull + i); }
uint64_t fake_mesh_bytes(uint64_t ntri) { return ntri * 40000; }

extern "C" {

const char *ib_last_error(void) { return g_err.c_str(); }

ib_status ib_create(int device, uint64_t, ib_context **out) {
  if (device != 0) return IB_ERR_NO_DEVICE;
  *out = new ib_context();
  return IB_OK;
}
void ib_destroy(ib_context *c) { delete c; }
ib_status ib_set_option(ib_context *c, int32_t opt, int64_t) {
  if (!c) return IB_ERR_INVALID_ARGUMENT;
  return opt == IB_OPT_SHARED_CHIP ? IB_OK : IB_ERR_INVALID_ARGUMENT;  // the only option the server sets
}

ib_status ib_mesh_create(ib_context *c, const float *d_xyz, uint64_t ntri, void *, ib_mesh **out) {
  std::lock_guard<std::mutex> lk(c->mu);
  auto *m = new ib_mesh{c, std::vector<float>(d_xyz, d_xyz + 9 * ntri)};
  *out = m;
  return IB_OK;
}
void ib_mesh_destroy(ib_mesh *m) {
  if (!m) return;
  std::lock_guard<std::mutex> lk(m->ctx->mu);  // the owning context's lock, as the library takes it
  m->soup.clear();
  delete m;
}
uint64_t ib_mesh_bytes(const ib_mesh *m) { return m ? fake_mesh_bytes(m->soup.size() / 9) : 0; }
uint64_t ib_mesh_triangles(const ib_mesh *m) { return m ? m->soup.size() / 9 : 0; }

ib_status ib_bake_batch_device(ib_context *c, ib_job *jobs, uint32_t njobs, void *) {
  std::lock_guard<std::mutex> lk(c->mu);
  std::this_thread::sleep_for(std::chrono::microseconds(200));
  for (uint32_t j = 0; j < njobs; ++j) {
    ib_job &jb = jobs[j];
    const float *soup = jb.mesh ? jb.mesh->soup.data() : jb.d_xyz;
    const uint64_t ntri = jb.mesh ? jb.mesh->soup.size() / 9 : jb.ntri;
    if (ntri == 13) {
      g_err = "job " + std::to_string(j) + ": the fake bake refuses 13 triangles";
      return IB_ERR_INVALID_MESH;
    }
    const uint64_t key = fake_soup_key(soup, ntri);
    if (jb.sdf.nx) {
      if (jb.band_voxels > 0) {
        const uint64_t nb = (uint64_t)(jb.sdf.nx / 8) * (jb.sdf.ny / 8) * (jb.sdf.nz / 8);
        uint64_t slot = 0;
        for (uint64_t b = 0; b < nb; ++b) jb.d_brick_index[b] = fake_expected_brick(key, b) ? (int32_t)slot++ : -1;
        jb.nbricks = slot;
        if (slot > jb.pool_bricks) {
          g_err = "pool too small";
          return IB_ERR_CAPACITY;
        }
        for (uint64_t i = 0; i < slot * 512; ++i) jb.d_distances[i] = fake_expected_distance(key, i);
      } else {
        const uint64_t nv = (uint64_t)jb.sdf.nx * jb.sdf.ny * jb.sdf.nz;
        for (uint64_t i = 0; i < nv; ++i) jb.d_distances[i] = fake_expected_distance(key, i);
      }
    }
    if (jb.occ.nx) {
      const uint64_t nw = (uint64_t)((jb.occ.nx + 31) / 32) * jb.occ.ny * jb.occ.nz;
      for (uint64_t i = 0; i < nw; ++i) jb.d_occ_words[i] = fake_expected_word(key, i);
    }
  }
  return IB_OK;
}

}  // extern "C"
