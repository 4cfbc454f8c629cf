// Drives the bake server (ib_server.cpp, built for the host over fake_device.cpp) from several client
// threads at once: dense, narrow-band and occupancy requests, keyed and unkeyed, under a budget that keeps
// requests queued and evicts resident meshes; every fetched result is checked against what the fake bake
// writes.  Then a request larger than the budget and one the bake refuses.  For the sanitizer builds
// (ThreadSanitizer; AddressSanitizer + UndefinedBehaviorSanitizer).  Exit status 0 = clean.
#include <atomic>
#include <cstdio>
#include <random>
#include <thread>
#include <vector>

#include "fake_device.h"
#include "italy_bake.h"

static std::atomic<int> failures{0};
#define CHECK(c)                                                        \
  do {                                                                  \
    if (!(c)) {                                                         \
      std::fprintf(stderr, "check failed at line %d: %s\n", __LINE__, #c); \
      ++failures;                                                       \
    }                                                                   \
  } while (0)

static ib_grid grid(int nx, int ny, int nz) {
  ib_grid g{};
  g.origin[0] = g.origin[1] = g.origin[2] = -1.0f;
  g.h = 2.0f / (float)nx;
  g.nx = nx;
  g.ny = ny;
  g.nz = nz;
  return g;
}

int main() {
  ib_server *s = nullptr;
  CHECK(ib_server_create(0, 96ull << 20, 3, 4, &s) == IB_OK);
  std::vector<std::vector<float>> soups(8);
  std::mt19937 gen(7);
  std::uniform_real_distribution<float> u(-0.9f, 0.9f);
  for (size_t m = 0; m < soups.size(); ++m) {
    soups[m].resize(9 * (100 + 100 * m));
    for (float &v : soups[m]) v = u(gen);
  }
  const int kClients = 4, kRequests = 30;
  auto client = [&](int c) {
    std::mt19937 rng(100 + c);
    for (int r = 0; r < kRequests; ++r) {
      const int m = (int)(rng() % soups.size());
      const std::vector<float> &soup = soups[m];
      ib_request q{};
      q.xyz = soup.data();
      q.ntri = soup.size() / 9;
      const int kind = (int)(rng() % 3);
      if (kind == 0) q.sdf = grid(16, 16, 16);
      if (kind == 1) {
        q.sdf = grid(32, 32, 32);
        q.band_voxels = 2;
      }
      if (kind == 2 || rng() % 2) q.occ = grid(40, 8, 8);
      q.mesh_key = (rng() % 2) ? 1000 + (uint64_t)m : 0;
      uint64_t t = 0;
      CHECK(ib_server_submit(s, &q, &t) == IB_OK);
      CHECK(ib_server_wait(s, t, 60.0) == IB_OK);
      ib_ticket_info info{};
      CHECK(ib_server_info(s, t, &info) == IB_OK && info.state == 2);
      const uint64_t key = fake_soup_key(soup.data(), q.ntri);
      const uint64_t nb = q.band_voxels ? (uint64_t)(q.sdf.nx / 8) * (q.sdf.ny / 8) * (q.sdf.nz / 8) : 0;
      const uint64_t nf = q.band_voxels ? info.nbricks * 512 : (q.sdf.nx ? (uint64_t)q.sdf.nx * q.sdf.ny * q.sdf.nz : 0);
      const uint64_t nw = q.occ.nx ? (uint64_t)((q.occ.nx + 31) / 32) * q.occ.ny * q.occ.nz : 0;
      std::vector<float> d(nf ? nf : 1);
      std::vector<int32_t> idx(nb ? nb : 1);
      std::vector<uint32_t> w(nw ? nw : 1);
      CHECK(ib_server_fetch(s, t, nf ? d.data() : nullptr, nf, nb ? idx.data() : nullptr, nw ? w.data() : nullptr) ==
            IB_OK);
      for (uint64_t i = 0; i < nf; ++i) CHECK(d[i] == fake_expected_distance(key, i));
      uint64_t slot = 0;
      for (uint64_t b = 0; b < nb; ++b) CHECK(idx[b] == (fake_expected_brick(key, b) ? (int32_t)slot++ : -1));
      if (nb) CHECK(slot == info.nbricks);
      for (uint64_t i = 0; i < nw; ++i) CHECK(w[i] == fake_expected_word(key, i));
      CHECK(ib_server_release(s, t) == IB_OK);
    }
  };
  std::vector<std::thread> threads;
  for (int c = 0; c < kClients; ++c) threads.emplace_back(client, c);
  for (auto &th : threads) th.join();

  // a request that can never fit the budget, and one the bake refuses
  ib_request big{};
  big.xyz = soups[0].data();
  big.ntri = 200000;
  big.sdf = grid(16, 16, 16);
  uint64_t tb = 0, tf = 0;
  CHECK(ib_server_submit(s, &big, &tb) == IB_OK);
  CHECK(ib_server_wait(s, tb, 10.0) == IB_ERR_OUT_OF_BUDGET);
  std::vector<float> bad(9 * 13, 0.25f);
  ib_request refused{};
  refused.xyz = bad.data();
  refused.ntri = 13;
  refused.sdf = grid(16, 16, 16);
  CHECK(ib_server_submit(s, &refused, &tf) == IB_OK);
  CHECK(ib_server_wait(s, tf, 10.0) == IB_ERR_INVALID_MESH);
  CHECK(ib_server_drain(s, 10.0) == IB_OK);
  CHECK(ib_server_release(s, tb) == IB_OK && ib_server_release(s, tf) == IB_OK);
  ib_server_counters st{};
  CHECK(ib_server_stats(s, &st) == IB_OK);
  CHECK(st.completed == (uint64_t)(kClients * kRequests) + 1);
  CHECK(st.failed == 2);
  CHECK(st.in_use == st.mesh_bytes);  // every result released: only resident meshes remain
  CHECK(st.mesh_builds > 0 && st.mesh_hits > 0 && st.evictions > 0);
  std::printf("server drive: %llu batches, %llu mesh builds, %llu hits, %llu evictions\n",
              (unsigned long long)st.batches, (unsigned long long)st.mesh_builds, (unsigned long long)st.mesh_hits,
              (unsigned long long)st.evictions);
  ib_server_destroy(s);
  if (failures) {
    std::fprintf(stderr, "%d checks failed\n", failures.load());
    return 1;
  }
  std::printf("server drive: ok\n");
  return 0;
}
