// Bake server: admits requests by an HBM budget, keeps results resident on the device until they
// are fetched, and bakes whatever is admitted in batches on several worker contexts at once
// (continuous batching: requests that arrive while batches run join the next free worker's batch;
// the workers' batches run concurrently on their own streams).  Built on the public C ABI.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <mutex>
#include <string>
#include <thread>
#include <unordered_map>
#include <vector>

#include "italy_bake.h"

namespace {

thread_local std::string g_srv_err;

ib_status sfail(ib_status s, const std::string &m) {
  g_srv_err = m;
  return s;
}

double now_s() {
  using namespace std::chrono;
  return duration<double>(steady_clock::now().time_since_epoch()).count();
}

// Device blocks recycled by size class (powers of two up to 64 MiB, 16 MiB steps above) so that a
// steady request stream does not call hipMalloc / hipFree per request.
struct DevCache {
  std::mutex mu;
  std::map<size_t, std::vector<void *>> free_blocks;
  std::atomic<size_t> cached{0};  // read by the server's admission without this cache's lock
  static size_t klass(size_t n) {
    if (n <= (64ull << 20)) {
      size_t c = 4096;
      while (c < n) c <<= 1;
      return c;
    }
    const size_t step = 16ull << 20;
    return (n + step - 1) / step * step;
  }
  std::atomic<uint64_t> failures{0};  // hipMalloc calls that failed even after the cache was returned (read unlocked)
  void *get(size_t n) {
    const size_t k = klass(std::max<size_t>(n, 1));
    {
      std::lock_guard<std::mutex> lk(mu);
      auto it = free_blocks.find(k);
      if (it != free_blocks.end() && !it->second.empty()) {
        void *p = it->second.back();
        it->second.pop_back();
        cached -= k;
        return p;
      }
    }
    void *p = nullptr;
    if (hipMalloc(&p, k) == hipSuccess) return p;
    clear();  // out of device memory: give the cached blocks back and try once more
    if (hipMalloc(&p, k) == hipSuccess) return p;
    ++failures;
    return nullptr;
  }
  // frees cached blocks, largest classes first, until at most `keep` bytes stay cached
  void trim(size_t keep) {
    std::lock_guard<std::mutex> lk(mu);
    while (cached > keep && !free_blocks.empty()) {
      auto it = std::prev(free_blocks.end());
      while (!it->second.empty() && cached > keep) {
        (void)hipFree(it->second.back());
        it->second.pop_back();
        cached -= it->first;
      }
      if (it->second.empty()) free_blocks.erase(it);
    }
  }
  void put(void *p, size_t n) {
    if (!p) return;
    const size_t k = klass(std::max<size_t>(n, 1));
    std::lock_guard<std::mutex> lk(mu);
    free_blocks[k].push_back(p);
    cached += k;
  }
  void clear() {
    std::lock_guard<std::mutex> lk(mu);
    for (auto &kv : free_blocks)
      for (void *p : kv.second) (void)hipFree(p);
    free_blocks.clear();
    cached = 0;
  }
};

enum { ST_QUEUED = 0, ST_RUNNING = 1, ST_DONE = 2, ST_FAILED = 3 };

struct Ticket {
  ib_request req{};
  int state = ST_QUEUED;
  ib_status status = IB_OK;
  std::string err;
  double t_submit = 0, t_admit = 0, t_start = 0, t_done = 0;
  uint64_t est = 0;        // bytes reserved against the budget while admitted
  uint64_t resident = 0;   // bytes kept after the bake (outputs)
  void *d_xyz = nullptr;
  float *d_out = nullptr;
  int32_t *d_index = nullptr;
  uint32_t *d_occ = nullptr;
  uint64_t out_bytes = 0, index_bytes = 0, occ_bytes = 0, pool_bricks = 0, nbricks = 0;
  int worker = -1;
  int fetching = 0;        // fetches copying the result right now; release waits for none
  const ib_mesh *mesh = nullptr;  // resident mesh used by the running bake (cache entry pinned)
};

// A resident mesh kept under the request's mesh_key; `users` pins it while bakes use it.
struct CachedMesh {
  ib_mesh *mesh = nullptr;
  uint64_t bytes = 0;
  uint64_t last_use = 0;
  int users = 0;
  int owner = 0;  // worker whose context holds the mesh's memory
};

uint64_t voxels(const ib_grid &g) { return (uint64_t)g.nx * g.ny * g.nz; }

}  // namespace

struct ib_server {
  int device = 0;
  uint64_t budget = 0;
  uint64_t in_use = 0;
  int max_batch = 64;
  std::mutex mu;
  std::condition_variable cv_work, cv_done;
  std::deque<uint64_t> pending, ready;
  std::unordered_map<uint64_t, Ticket> tickets;
  uint64_t next_id = 1;
  bool stop = false;
  int running = 0;
  std::vector<std::thread> workers;
  std::vector<ib_context *> ctxs;
  DevCache cache;
  uint64_t batches = 0, completed = 0;
  std::unordered_map<uint64_t, CachedMesh> meshes;  // mesh_key -> resident mesh
  uint64_t mesh_clock = 0, mesh_hits = 0, mesh_builds = 0, mesh_bytes = 0, evictions = 0, failed = 0;
  // evicted meshes waiting for their owning worker to destroy them between its batches (destroying one
  // takes the owner's context lock, which the owner holds for a whole batch)
  std::vector<std::vector<ib_mesh *>> doomed;
  uint64_t doomed_bytes = 0;

  // device bytes in use or held: admitted requests and results, resident meshes, meshes awaiting their
  // destruction and the staging cache
  uint64_t held_locked() const { return in_use + doomed_bytes + cache.cached; }

  // makes room for `need` more bytes: trims the staging cache, then evicts the least recently used
  // resident meshes that no bake is using (their owners destroy them)
  void evict_locked(uint64_t need) {
    if (held_locked() + need > budget && cache.cached) {
      const uint64_t over = held_locked() + need - budget;
      cache.trim(cache.cached > over ? cache.cached - over : 0);
    }
    while (in_use + doomed_bytes + need > budget) {
      auto victim = meshes.end();
      for (auto it = meshes.begin(); it != meshes.end(); ++it)
        if (it->second.users == 0 && (victim == meshes.end() || it->second.last_use < victim->second.last_use))
          victim = it;
      if (victim == meshes.end()) return;
      in_use -= victim->second.bytes;
      mesh_bytes -= victim->second.bytes;
      doomed_bytes += victim->second.bytes;
      doomed[victim->second.owner].push_back(victim->second.mesh);
      ++evictions;
      meshes.erase(victim);
      cv_work.notify_all();
    }
  }

  // destroys the evicted meshes worker w owns; called by w between batches, without the server lock
  void reap(int w) {
    std::vector<ib_mesh *> mine;
    {
      std::lock_guard<std::mutex> lk(mu);
      mine.swap(doomed[w]);
    }
    if (mine.empty()) return;
    uint64_t freed = 0;
    for (ib_mesh *m : mine) {
      freed += ib_mesh_bytes(m);
      ib_mesh_destroy(m);
    }
    std::lock_guard<std::mutex> lk(mu);
    doomed_bytes -= freed;
    try_admit_locked();
  }

  // device bytes a request needs while it is baked, and what stays resident afterwards
  static void estimate(const ib_request &r, uint64_t &est, uint64_t &resident) {
    const uint64_t n = r.ntri;
    uint64_t out = 0;
    if (r.sdf.nx) {
      if (r.band_voxels > 0) {
        const uint64_t bricks = (uint64_t)(r.sdf.nx / 8) * (r.sdf.ny / 8) * (r.sdf.nz / 8);
        out = bricks * 2048 + bricks * 4;
      } else {
        out = voxels(r.sdf) * 4;
      }
    }
    if (r.occ.nx) out += (uint64_t)((r.occ.nx + 31) / 32) * r.occ.ny * r.occ.nz * 4;
    const uint64_t cols = r.sdf.nx ? (uint64_t)r.sdf.nx * r.sdf.ny * (32 * 8 + 4) : 0;
    // soup + triangle and distance records + nodes (binary and 4-wide) + oriented leaf boxes + weld/edge tables
    // + sort temporaries + refit leaf boxes + columns + lists
    est = n * (36 + 48 + 96 + 64 + 128 + 128 + 320 + 24) + cols + out / 8 + out + (8u << 20);
    resident = out;
  }

  void try_admit_locked() {
    while (!pending.empty()) {
      Ticket &t = tickets[pending.front()];
      if (held_locked() + t.est > budget) evict_locked(t.est);
      if (held_locked() + t.est > budget) {
        if (t.est > budget) {  // can never fit
          t.state = ST_FAILED;
          t.status = IB_ERR_OUT_OF_BUDGET;
          t.err = "request needs " + std::to_string(t.est) + " bytes, the budget is " + std::to_string(budget);
          t.t_done = now_s();
          ++failed;
          pending.pop_front();
          cv_done.notify_all();
          continue;
        }
        break;  // FIFO admission: wait for memory to free up
      }
      in_use += t.est;
      t.t_admit = now_s();
      ready.push_back(pending.front());
      pending.pop_front();
      cv_work.notify_one();
    }
  }

  void worker_loop(int w) {
    ib_context *ctx = ctxs[w];
    (void)hipSetDevice(device);
    std::vector<uint64_t> batch;
    std::vector<ib_job> jobs;
    for (;;) {
      batch.clear();
      reap(w);
      {
        std::unique_lock<std::mutex> lk(mu);
        cv_work.wait(lk, [&] { return stop || !ready.empty() || !doomed[w].empty(); });
        if (!doomed[w].empty() && !stop) continue;  // destroy evicted meshes first, outside the lock
        if (stop && ready.empty()) return;
        while (!ready.empty() && (int)batch.size() < max_batch) {
          batch.push_back(ready.front());
          ready.pop_front();
        }
        ++running;
      }
      // stage the soups and outputs (outside the lock)
      jobs.assign(batch.size(), ib_job{});
      const double t0 = now_s();
      for (size_t k = 0; k < batch.size(); ++k) {
        Ticket *t;
        {
          std::lock_guard<std::mutex> lk(mu);
          t = &tickets[batch[k]];
          t->state = ST_RUNNING;
          t->t_start = t0;
          t->worker = w;
        }
        const ib_request &r = t->req;
        ib_job &j = jobs[k];
        j.ntri = r.ntri;
        bool ok = true;
        if (r.mesh_key && r.ntri) {  // a resident mesh under this key: no upload, no build
          std::lock_guard<std::mutex> lk(mu);
          auto it = meshes.find(r.mesh_key);
          if (it != meshes.end() && ib_mesh_triangles(it->second.mesh) == r.ntri) {
            ++it->second.users;
            it->second.last_use = ++mesh_clock;
            ++mesh_hits;
            t->mesh = it->second.mesh;
            j.mesh = t->mesh;
          }
        }
        if (r.ntri && !j.mesh) {
          t->d_xyz = cache.get(r.ntri * 36);
          ok &= t->d_xyz && hipMemcpy(t->d_xyz, r.xyz, r.ntri * 36, hipMemcpyHostToDevice) == hipSuccess;
          j.d_xyz = static_cast<const float *>(t->d_xyz);
        }
        if (r.sdf.nx) {
          j.sdf = r.sdf;
          if (r.band_voxels > 0) {
            t->pool_bricks = (uint64_t)(r.sdf.nx / 8) * (r.sdf.ny / 8) * (r.sdf.nz / 8);
            t->out_bytes = t->pool_bricks * 2048;
            t->index_bytes = t->pool_bricks * 4;
            t->d_index = static_cast<int32_t *>(cache.get(t->index_bytes));
            j.band_voxels = r.band_voxels;
            j.d_brick_index = t->d_index;
            j.pool_bricks = t->pool_bricks;
            ok &= t->d_index != nullptr;
          } else {
            t->out_bytes = voxels(r.sdf) * 4;
          }
          t->d_out = static_cast<float *>(cache.get(t->out_bytes));
          j.d_distances = t->d_out;
          ok &= t->d_out != nullptr;
        }
        if (r.occ.nx) {
          t->occ_bytes = (uint64_t)((r.occ.nx + 31) / 32) * r.occ.ny * r.occ.nz * 4;
          t->d_occ = static_cast<uint32_t *>(cache.get(t->occ_bytes));
          j.occ = r.occ;
          j.d_occ_words = t->d_occ;
          ok &= t->d_occ != nullptr;
        }
        if (!ok) {
          std::lock_guard<std::mutex> lk(mu);
          t->status = IB_ERR_DEVICE;
          t->err = "staging the request on the device failed (hipMalloc or the soup upload)";
        }
      }
      // keyed requests that missed: build their mesh once, keep it resident under the key
      for (size_t k = 0; k < batch.size(); ++k) {
        Ticket *t;
        {
          std::lock_guard<std::mutex> lk(mu);
          t = &tickets[batch[k]];
        }
        if (!t->req.mesh_key || t->mesh || !t->d_xyz || t->status != IB_OK) continue;
        // use the entry if it appeared since staging (an earlier job of this batch, another worker); a key
        // whose entry holds a different triangle count is baked from its soup and not cached
        const auto take_existing = [&]() -> bool {
          auto it = meshes.find(t->req.mesh_key);
          if (it == meshes.end()) return false;
          if (ib_mesh_triangles(it->second.mesh) == t->req.ntri) {
            ++it->second.users;
            it->second.last_use = ++mesh_clock;
            ++mesh_hits;
            t->mesh = it->second.mesh;
            jobs[k].mesh = t->mesh;
          }
          return true;
        };
        {
          std::lock_guard<std::mutex> lk(mu);
          if (take_existing()) continue;
        }
        ib_mesh *mh = nullptr;
        if (ib_mesh_create(ctx, static_cast<const float *>(t->d_xyz), t->req.ntri, nullptr, &mh) != IB_OK) continue;
        std::lock_guard<std::mutex> lk(mu);
        if (take_existing()) {
          ib_mesh_destroy(mh);
          continue;
        }
        {
          CachedMesh cm;
          cm.mesh = mh;
          cm.owner = w;
          cm.bytes = ib_mesh_bytes(mh);
          cm.last_use = ++mesh_clock;
          cm.users = 1;
          meshes.emplace(t->req.mesh_key, cm);
          in_use += cm.bytes;
          mesh_bytes += cm.bytes;
          ++mesh_builds;
          t->mesh = mh;
        }
        jobs[k].mesh = t->mesh;
      }
      // drop requests that could not be staged, bake the rest as one batch
      std::vector<ib_job> run;
      std::vector<uint64_t> ids;
      {
        std::lock_guard<std::mutex> lk(mu);
        for (size_t k = 0; k < batch.size(); ++k)
          if (tickets[batch[k]].status == IB_OK) {
            run.push_back(jobs[k]);
            ids.push_back(batch[k]);
          }
      }
      const ib_status st = run.empty() ? IB_OK : ib_bake_batch_device(ctx, run.data(), (uint32_t)run.size(), nullptr);
      const std::string err = st == IB_OK ? "" : ib_last_error();
      const double t1 = now_s();
      {
        std::lock_guard<std::mutex> lk(mu);
        for (size_t k = 0; k < batch.size(); ++k) {
          Ticket &t = tickets[batch[k]];
          if (t.status == IB_OK && st != IB_OK) {
            t.status = st;
            t.err = err;
          }
          for (size_t q = 0; q < ids.size(); ++q)
            if (ids[q] == batch[k]) t.nbricks = run[q].nbricks;
          t.state = t.status == IB_OK ? ST_DONE : ST_FAILED;
          t.t_done = t1;
          cache.put(t.d_xyz, t.req.ntri * 36);  // the soup is no longer needed on the device
          t.d_xyz = nullptr;
          if (t.mesh) {
            auto it = meshes.find(t.req.mesh_key);
            if (it != meshes.end() && it->second.mesh == t.mesh) --it->second.users;
            t.mesh = nullptr;
          }
          t.resident = t.out_bytes + t.index_bytes + t.occ_bytes;
          in_use -= t.est;
          in_use += t.resident;
          if (t.state == ST_FAILED) {  // failed requests keep nothing
            ++failed;
            cache.put(t.d_out, t.out_bytes);
            cache.put(t.d_index, t.index_bytes);
            cache.put(t.d_occ, t.occ_bytes);
            t.d_out = nullptr;
            t.d_index = nullptr;
            t.d_occ = nullptr;
            in_use -= t.resident;
            t.resident = 0;
          }
          ++completed;
        }
        ++batches;
        --running;
        try_admit_locked();
      }
      cv_done.notify_all();
    }
  }
};

extern "C" {

const char *ib_server_last_error(void) { return g_srv_err.c_str(); }

ib_status ib_server_create(int device, uint64_t hbm_budget, int32_t workers, int32_t max_batch, ib_server **out) {
  if (!out || workers < 1 || workers > 16 || max_batch < 1) return sfail(IB_ERR_INVALID_ARGUMENT, "bad arguments");
  *out = nullptr;
  auto *s = new ib_server();
  s->device = device;
  s->max_batch = max_batch;
  if (hipSetDevice(device) != hipSuccess) {
    delete s;
    return sfail(IB_ERR_NO_DEVICE, "no such device");
  }
  size_t freeb = 0, total = 0;
  (void)hipMemGetInfo(&freeb, &total);
  s->budget = hbm_budget ? hbm_budget : (uint64_t)(0.8 * (double)freeb);
  s->doomed.resize(workers);
  for (int w = 0; w < workers; ++w) {
    // every worker context gets the whole budget (not 90% of free memory each): the server's admission
    // is what divides the budget between the workers' batches
    ib_context *c = nullptr;
    ib_status st = ib_create(device, s->budget, &c);
    if (st == IB_OK) st = ib_set_option(c, IB_OPT_SHARED_CHIP, 1);  // the workers' batches run side by side
    if (st != IB_OK) {
      if (c) ib_destroy(c);
      for (auto *x : s->ctxs) ib_destroy(x);
      delete s;
      return sfail(st, ib_last_error());
    }
    s->ctxs.push_back(c);
  }
  for (int w = 0; w < workers; ++w) s->workers.emplace_back([s, w] { s->worker_loop(w); });
  *out = s;
  return IB_OK;
}

void ib_server_destroy(ib_server *s) {
  if (!s) return;
  {
    std::lock_guard<std::mutex> lk(s->mu);
    s->stop = true;
  }
  s->cv_work.notify_all();
  for (auto &t : s->workers) t.join();
  for (auto &kv : s->tickets) {
    Ticket &t = kv.second;
    s->cache.put(t.d_out, t.out_bytes);
    s->cache.put(t.d_index, t.index_bytes);
    s->cache.put(t.d_occ, t.occ_bytes);
    s->cache.put(t.d_xyz, t.req.ntri * 36);
  }
  for (auto &kv : s->meshes) ib_mesh_destroy(kv.second.mesh);  // before the contexts that own their memory
  s->meshes.clear();
  for (auto &v : s->doomed)
    for (ib_mesh *m : v) ib_mesh_destroy(m);
  s->doomed.clear();
  s->cache.clear();
  for (auto *c : s->ctxs) ib_destroy(c);
  delete s;
}

ib_status ib_server_submit(ib_server *s, const ib_request *r, uint64_t *ticket) {
  if (!s || !r || !ticket) return sfail(IB_ERR_INVALID_ARGUMENT, "null argument");
  if (r->ntri && !r->xyz) return sfail(IB_ERR_INVALID_ARGUMENT, "null triangle buffer");
  if (r->band_voxels > 0 && (!r->sdf.nx || (r->sdf.nx | r->sdf.ny | r->sdf.nz) & 7))
    return sfail(IB_ERR_INVALID_ARGUMENT, "a narrow band needs grid dims divisible by 8");
  Ticket t;
  t.req = *r;
  t.t_submit = now_s();
  uint64_t out_estimate = 0;  // what the result will hold; counted (as t.resident) only once it exists
  ib_server::estimate(*r, t.est, out_estimate);
  std::lock_guard<std::mutex> lk(s->mu);
  const uint64_t id = s->next_id++;
  s->tickets.emplace(id, t);
  s->pending.push_back(id);
  s->try_admit_locked();
  *ticket = id;
  return IB_OK;
}

ib_status ib_server_info(ib_server *s, uint64_t ticket, ib_ticket_info *info) {
  if (!s || !info) return sfail(IB_ERR_INVALID_ARGUMENT, "null argument");
  std::lock_guard<std::mutex> lk(s->mu);
  auto it = s->tickets.find(ticket);
  if (it == s->tickets.end()) return sfail(IB_ERR_INVALID_ARGUMENT, "unknown ticket");
  const Ticket &t = it->second;
  info->state = t.state;
  info->status = t.status;
  info->t_submit = t.t_submit;
  info->t_admit = t.t_admit;
  info->t_start = t.t_start;
  info->t_done = t.t_done;
  info->nbricks = t.nbricks;
  info->resident_bytes = t.resident;
  info->worker = t.worker;
  return IB_OK;
}

ib_status ib_server_wait(ib_server *s, uint64_t ticket, double timeout_s) {
  if (!s) return sfail(IB_ERR_INVALID_ARGUMENT, "null server");
  std::unique_lock<std::mutex> lk(s->mu);
  auto it = s->tickets.find(ticket);
  if (it == s->tickets.end()) return sfail(IB_ERR_INVALID_ARGUMENT, "unknown ticket");
  const auto done = [&] { return it->second.state == ST_DONE || it->second.state == ST_FAILED; };
  if (!s->cv_done.wait_for(lk, std::chrono::duration<double>(timeout_s), done))
    return sfail(IB_ERR_NOT_READY, "ticket still running");
  if (it->second.state == ST_FAILED) return sfail(it->second.status, it->second.err);
  return IB_OK;
}

ib_status ib_server_drain(ib_server *s, double timeout_s) {
  if (!s) return sfail(IB_ERR_INVALID_ARGUMENT, "null server");
  std::unique_lock<std::mutex> lk(s->mu);
  const bool ok = s->cv_done.wait_for(lk, std::chrono::duration<double>(timeout_s), [&] {
    return s->pending.empty() && s->ready.empty() && s->running == 0;
  });
  return ok ? IB_OK : sfail(IB_ERR_NOT_READY, "work still queued");
}

ib_status ib_server_fetch(ib_server *s, uint64_t ticket, float *distances, uint64_t capacity, int32_t *brick_index,
                          uint32_t *occ_words) {
  if (!s) return sfail(IB_ERR_INVALID_ARGUMENT, "null server");
  Ticket t;
  {
    std::lock_guard<std::mutex> lk(s->mu);
    auto it = s->tickets.find(ticket);
    if (it == s->tickets.end()) return sfail(IB_ERR_INVALID_ARGUMENT, "unknown ticket");
    if (it->second.state == ST_FAILED) return sfail(it->second.status, it->second.err);
    if (it->second.state != ST_DONE) return sfail(IB_ERR_NOT_READY, "ticket still running");
    t = it->second;
    ++it->second.fetching;  // the device buffers stay owned by this ticket until the copies end
  }
  const auto copies = [&]() -> ib_status {
    (void)hipSetDevice(s->device);
    const uint64_t floats = t.req.band_voxels > 0 ? t.nbricks * 512 : (t.req.sdf.nx ? voxels(t.req.sdf) : 0);
    if (distances && floats) {
      if (capacity < floats) return sfail(IB_ERR_CAPACITY, "distance buffer holds " + std::to_string(capacity) +
                                                               " floats, the result has " + std::to_string(floats));
      if (hipMemcpy(distances, t.d_out, floats * 4, hipMemcpyDeviceToHost) != hipSuccess)
        return sfail(IB_ERR_DEVICE, "copying the distances failed");
    }
    if (brick_index && t.d_index &&
        hipMemcpy(brick_index, t.d_index, t.index_bytes, hipMemcpyDeviceToHost) != hipSuccess)
      return sfail(IB_ERR_DEVICE, "copying the brick index failed");
    if (occ_words && t.d_occ && hipMemcpy(occ_words, t.d_occ, t.occ_bytes, hipMemcpyDeviceToHost) != hipSuccess)
      return sfail(IB_ERR_DEVICE, "copying the occupancy failed");
    return IB_OK;
  };
  const ib_status st = copies();
  {
    std::lock_guard<std::mutex> lk(s->mu);
    --s->tickets[ticket].fetching;
  }
  s->cv_done.notify_all();
  return st;
}

ib_status ib_server_device_result(ib_server *s, uint64_t ticket, const float **d_distances, const int32_t **d_index,
                                  const uint32_t **d_occ) {
  if (!s) return sfail(IB_ERR_INVALID_ARGUMENT, "null server");
  std::lock_guard<std::mutex> lk(s->mu);
  auto it = s->tickets.find(ticket);
  if (it == s->tickets.end()) return sfail(IB_ERR_INVALID_ARGUMENT, "unknown ticket");
  if (it->second.state != ST_DONE) return sfail(IB_ERR_NOT_READY, "ticket not done");
  if (d_distances) *d_distances = it->second.d_out;
  if (d_index) *d_index = it->second.d_index;
  if (d_occ) *d_occ = it->second.d_occ;
  return IB_OK;
}

ib_status ib_server_release(ib_server *s, uint64_t ticket) {
  if (!s) return sfail(IB_ERR_INVALID_ARGUMENT, "null server");
  std::lock_guard<std::mutex> lk(s->mu);
  auto it = s->tickets.find(ticket);
  if (it == s->tickets.end()) return sfail(IB_ERR_INVALID_ARGUMENT, "unknown ticket");
  Ticket &t = it->second;
  if (t.state == ST_QUEUED || t.state == ST_RUNNING) return sfail(IB_ERR_NOT_READY, "ticket still running");
  if (t.fetching) return sfail(IB_ERR_NOT_READY, "ticket is being fetched");
  s->cache.put(t.d_out, t.out_bytes);
  s->cache.put(t.d_index, t.index_bytes);
  s->cache.put(t.d_occ, t.occ_bytes);
  s->in_use -= t.resident;
  s->tickets.erase(it);
  s->try_admit_locked();
  return IB_OK;
}

ib_status ib_server_stats(ib_server *s, ib_server_counters *out) {
  if (!s || !out) return sfail(IB_ERR_INVALID_ARGUMENT, "null argument");
  std::lock_guard<std::mutex> lk(s->mu);
  out->budget = s->budget;
  out->in_use = s->in_use;
  out->cached = s->cache.cached;
  out->pending = s->pending.size();
  out->ready = s->ready.size();
  out->running_batches = (uint64_t)s->running;
  out->batches = s->batches;
  out->completed = s->completed;
  out->mesh_hits = s->mesh_hits;
  out->mesh_builds = s->mesh_builds;
  out->meshes_resident = s->meshes.size();
  out->mesh_bytes = s->mesh_bytes;
  out->evictions = s->evictions;
  out->failed = s->failed;
  out->alloc_failures = s->cache.failures;
  return IB_OK;
}

}  // extern "C"
