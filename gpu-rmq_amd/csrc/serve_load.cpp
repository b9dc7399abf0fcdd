// serve_load.cpp — open-loop load generator for the serving loop (include/rmq_serve.h).
//
//   serve_load --arrays n:t:variant[,n:t:variant...] --rates r1[,r2...] [--seconds S] [--seed S]
//              [--slots K] [--max-batch-queries Q]
//
// Admits one resident array per spec (device memory filled with rmq_fill_uniform(seed + a)),
// then for every offered rate (requests/s) replays a Poisson arrival process over `seconds` of
// requests drawn from a fixed pool: target array by Zipf(1.1) over the arrays in spec order,
// 1..2^16 queries (log-uniform), lengths small / medium / large (one class per request; the
// lognormal scales of the reference generator, n^0.3 and n^0.6 with sigma 0.3, truncated and
// redrawn outside [1, n], or U_int[1, n]), l uniform in [0, n - L]. Index answers land in a ring
// of host memory. The client is one thread that submits every request whose scheduled arrival
// has passed and polls the server in between; latency = the time poll reports a request done -
// its scheduled arrival, so a client that falls behind shows up in the latency and in
// `client_lag` (a submit refused with RMQ_EAGAIN, the staging ring being full, is retried after
// the next poll). The client's time is split into submit, poll and the rest (the loop itself).
// One JSON line per rate on stdout (after a header line); a final line checks 32 requests
// answered by the server against GPU BASE (rmq_gpu_base) on the same device arrays.

#include <hip/hip_runtime_api.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <random>
#include <string>
#include <vector>

#include "../../include/rmq.h"
#include "../../include/rmq_batch.h"
#include "../../include/rmq_serve.h"

namespace {

using Clock = std::chrono::steady_clock;

struct Spec {
  uint64_t n;
  uint32_t t;
  uint32_t variant;
};

uint32_t variant_of(const std::string& s) {
  if (s == "lane") return RMQ_VARIANT_LANE;
  if (s == "lane8") return RMQ_VARIANT_LANE8;
  if (s == "lane4") return RMQ_VARIANT_LANE4;
  if (s == "group4") return RMQ_VARIANT_GROUP4;
  if (s == "group8") return RMQ_VARIANT_GROUP8;
  if (s == "group16") return RMQ_VARIANT_GROUP16;
  if (s == "group32") return RMQ_VARIANT_GROUP32;
  if (s == "wave") return RMQ_VARIANT_WAVE;
  fprintf(stderr, "unknown variant %s\n", s.c_str());
  exit(2);
}

std::vector<std::string> split(const std::string& s, char c) {
  std::vector<std::string> out;
  size_t a = 0;
  while (a <= s.size()) {
    const size_t b = s.find(c, a);
    out.push_back(s.substr(a, b == std::string::npos ? std::string::npos : b - a));
    if (b == std::string::npos) break;
    a = b + 1;
  }
  return out;
}

void die(const char* what, rmq_status st) {
  fprintf(stderr, "serve_load: %s failed: %s\n", what, rmq_status_str(st));
  exit(1);
}

// One request of the pool: its array and its {l, r} pairs at pairs[off .. off + 2q).
struct Req {
  uint32_t array;
  uint32_t q;
  uint64_t off;
};

// Length of one query of class k (0 small, 1 medium, 2 large) on an array of n elements.
uint64_t draw_length(std::mt19937_64& g, uint64_t n, int k) {
  if (k == 2) return std::uniform_int_distribution<uint64_t>(1, n)(g);
  const double scale = std::floor(std::pow(static_cast<double>(n), k == 0 ? 0.3 : 0.6));
  std::lognormal_distribution<double> d(std::log(std::max(1.0, scale)), 0.3);
  for (;;) {
    const double L = std::floor(d(g));
    if (L >= 1.0 && L <= static_cast<double>(n)) return static_cast<uint64_t>(L);
  }
}

double pct(std::vector<double> v, double p) {
  if (v.empty()) return 0.0;
  std::sort(v.begin(), v.end());
  const size_t i = std::min(v.size() - 1, static_cast<size_t>(p / 100.0 * (v.size() - 1) + 0.5));
  return v[i];
}

}  // namespace

int main(int argc, char** argv) {
  std::vector<Spec> specs;
  std::vector<double> rates;
  double seconds = 1.0;
  uint64_t seed = 27722;
  uint32_t slots = 3, max_batch = 1u << 21;
  for (int i = 1; i + 1 < argc; i += 2) {
    const std::string k = argv[i], v = argv[i + 1];
    if (k == "--arrays") {
      for (const std::string& x : split(v, ',')) {
        const std::vector<std::string> f = split(x, ':');
        if (f.size() != 3) {
          fprintf(stderr, "bad array spec %s\n", x.c_str());
          return 2;
        }
        specs.push_back(Spec{std::stoull(f[0]), static_cast<uint32_t>(std::stoul(f[1])), variant_of(f[2])});
      }
    } else if (k == "--rates") {
      for (const std::string& x : split(v, ',')) rates.push_back(std::stod(x));
    } else if (k == "--seconds") {
      seconds = std::stod(v);
    } else if (k == "--seed") {
      seed = std::stoull(v);
    } else if (k == "--slots") {
      slots = static_cast<uint32_t>(std::stoul(v));
    } else if (k == "--max-batch-queries") {
      max_batch = static_cast<uint32_t>(std::stoul(v));
    } else {
      fprintf(stderr, "unknown option %s\n", k.c_str());
      return 2;
    }
  }
  if (specs.empty() || rates.empty()) {
    fprintf(stderr, "usage: serve_load --arrays n:t:variant,... --rates r1,... [--seconds S]\n");
    return 2;
  }
  if (hipSetDevice(0) != hipSuccess) {
    fprintf(stderr, "serve_load: no HIP device\n");
    return 1;
  }

  rmq_server_options o{};
  o.slots = slots;
  o.max_batch_queries = max_batch;
  rmq_server* srv = nullptr;
  rmq_status st = rmq_server_create(&o, &srv);
  if (st != RMQ_OK) die("rmq_server_create", st);
  const uint32_t A = static_cast<uint32_t>(specs.size());
  std::vector<float*> arrays(A, nullptr);
  std::vector<int32_t> ids(A, -1);
  uint64_t array_bytes = 0;
  for (uint32_t a = 0; a < A; ++a) {
    if (hipMalloc(reinterpret_cast<void**>(&arrays[a]), specs[a].n * 4) != hipSuccess) die("hipMalloc", RMQ_ENOMEM);
    st = rmq_fill_uniform(arrays[a], specs[a].n, seed + a, 0, nullptr);
    if (st != RMQ_OK) die("rmq_fill_uniform", st);
    st = rmq_server_admit(srv, RMQ_ADMIT_DEVICE, arrays[a], specs[a].n, 0, specs[a].t, specs[a].variant, &ids[a]);
    if (st != RMQ_OK) die("rmq_server_admit", st);
    array_bytes += specs[a].n * 4;
  }
  rmq_server_stats ss{};
  rmq_server_get_stats(srv, &ss);
  printf("{\"arrays\": %u, \"resident_gb\": %.3f, \"array_gb\": %.3f, \"level_gb\": %.3f}\n", A,
         (array_bytes + ss.resident_bytes) / 1e9, array_bytes / 1e9, ss.resident_bytes / 1e9);
  fflush(stdout);

  // request pool
  std::mt19937_64 g(seed);
  std::vector<double> cdf(A);
  double z = 0;
  for (uint32_t a = 0; a < A; ++a) cdf[a] = (z += std::pow(a + 1.0, -1.1));
  for (double& c : cdf) c /= z;
  std::uniform_real_distribution<double> U(0.0, 1.0);
  const uint32_t kPool = 4096;
  std::vector<Req> pool(kPool);
  std::vector<uint32_t> pairs;
  for (uint32_t i = 0; i < kPool; ++i) {
    const uint32_t a = static_cast<uint32_t>(std::lower_bound(cdf.begin(), cdf.end(), U(g)) - cdf.begin());
    const uint32_t q = std::max<uint32_t>(1, static_cast<uint32_t>(std::exp2(U(g) * 16.0)));
    const int kind = static_cast<int>(U(g) * 3.0) % 3;
    const uint64_t n = specs[std::min(a, A - 1)].n;
    pool[i] = Req{std::min(a, A - 1), q, pairs.size()};
    for (uint32_t j = 0; j < q; ++j) {
      const uint64_t L = draw_length(g, n, kind);
      const uint64_t l = std::uniform_int_distribution<uint64_t>(0, n - L)(g);
      pairs.push_back(static_cast<uint32_t>(l));
      pairs.push_back(static_cast<uint32_t>(l + L - 1));
    }
  }

  // answer ring: request i's answers at ring[start % kRing ...] (a request never wraps)
  const uint64_t kRing = 1ull << 27;
  std::vector<uint32_t> ring(kRing);
  std::vector<uint64_t> done_buf(1 << 16);
  // warm-up, untimed: the first launches load the kernels' code and touch the staging rings, a
  // one-time cost that would otherwise land on the first rate's tail latencies
  const uint32_t kWarm = 256;
  for (uint32_t i = 0; i < kWarm; ++i) {
    const Req& r = pool[i];
    uint64_t tk = 0;
    uint32_t nd = 0;
    st = rmq_server_submit(srv, ids[r.array], pairs.data() + r.off, r.q, nullptr, ring.data(), &tk);
    if (st == RMQ_OK) st = rmq_server_drain(srv, done_buf.data(), static_cast<uint32_t>(done_buf.size()), &nd);
    if (st != RMQ_OK) die("warm-up request", st);
  }
  printf("{\"warmup_requests\": %u}\n", kWarm);
  fflush(stdout);
  for (double rate : rates) {
    const uint64_t N = std::max<uint64_t>(64, static_cast<uint64_t>(rate * seconds));
    std::mt19937_64 ga(seed ^ static_cast<uint64_t>(rate));
    std::exponential_distribution<double> gap(rate);
    std::vector<double> arrive(N);
    double t = 0;
    for (uint64_t i = 0; i < N; ++i) arrive[i] = (t += gap(ga));
    std::vector<uint64_t> start(N), ticket_req;
    std::vector<double> finish(N, -1.0), lag(N, 0.0);
    std::vector<uint8_t> done(N, 0);
    uint64_t total_q = 0;
    for (uint64_t i = 0; i < N; ++i) total_q += pool[i % kPool].q;
    rmq_server_get_stats(srv, &ss);
    const uint64_t batches0 = ss.batches, first_ticket = ss.submitted + 1;
    uint64_t next = 0, oldest = 0, alloc = 0, completed = 0, refused = 0;
    double t_submit = 0, t_poll = 0;
    const Clock::time_point t0 = Clock::now();
    auto now = [&]() { return std::chrono::duration<double>(Clock::now() - t0).count(); };
    while (completed < N) {
      const double tn = now();
      while (next < N && arrive[next] <= tn) {
        const Req& r = pool[next % kPool];
        uint64_t s0 = alloc;
        if (s0 % kRing + r.q > kRing) s0 += kRing - s0 % kRing;  // no wrap inside a request
        while (oldest < next && done[oldest]) ++oldest;
        if (oldest < next && s0 + r.q - start[oldest] > kRing) break;  // ring full: poll first
        uint64_t tk = 0;
        const double ts = now();
        st = rmq_server_submit(srv, ids[r.array], pairs.data() + r.off, r.q, nullptr, ring.data() + s0 % kRing, &tk);
        t_submit += now() - ts;
        if (st == RMQ_EAGAIN) {  // the server's staging is full: poll first
          ++refused;
          break;
        }
        if (st != RMQ_OK) die("rmq_server_submit", st);
        start[next] = s0;
        alloc = s0 + r.q;
        lag[next] = now() - arrive[next];
        ++next;
      }
      uint32_t nd = 0;
      const double tp = now();
      st = rmq_server_poll(srv, done_buf.data(), static_cast<uint32_t>(done_buf.size()), &nd);
      t_poll += now() - tp;
      if (st != RMQ_OK) die("rmq_server_poll", st);
      if (nd) {
        const double tf = now();
        for (uint32_t k = 0; k < nd; ++k) {
          const uint64_t i = done_buf[k] - first_ticket;  // tickets are issued in submission order
          done[i] = 1;
          finish[i] = tf;
          ++completed;
        }
      }
    }
    std::vector<double> latency(N);
    double last = 0;
    for (uint64_t i = 0; i < N; ++i) {
      latency[i] = (finish[i] - arrive[i]) * 1e3;
      last = std::max(last, finish[i]);
    }
    rmq_server_get_stats(srv, &ss);
    const uint64_t batches = ss.batches - batches0;
    const double span = last - arrive[0];
    const double offered = total_q / (arrive[N - 1] - arrive[0] + 1e-12) / 1e9;
    const double achieved = total_q / span / 1e9;
    const double lag99 = pct(lag, 99) * 1e3;
    printf("{\"offered_rps\": %.0f, \"requests\": %llu, \"queries\": %llu, \"offered_gqs\": %.4f, "
           "\"gqs\": %.4f, \"p50_ms\": %.3f, \"p90_ms\": %.3f, \"p99_ms\": %.3f, \"max_ms\": %.3f, "
           "\"batches\": %llu, \"requests_per_batch\": %.1f, \"client_lag_p99_ms\": %.3f, "
           "\"client_submit_s\": %.3f, \"client_poll_s\": %.3f, \"wall_s\": %.3f, \"refused\": %llu, "
           "\"saturated\": %s}\n",
           rate, (unsigned long long)N, (unsigned long long)total_q, offered, achieved, pct(latency, 50),
           pct(latency, 90), pct(latency, 99), pct(latency, 100), (unsigned long long)batches,
           batches ? static_cast<double>(N) / batches : 0.0, lag99, t_submit, t_poll, last,
           (unsigned long long)refused,
           (achieved < 0.9 * offered || lag99 > 1.0) ? "true" : "false");
    fflush(stdout);
  }

  // 32 small / medium requests of the pool against GPU BASE on the same device arrays
  uint32_t checked = 0, wrong = 0;
  for (uint32_t i = 0; i < kPool && checked < 32; ++i) {
    const Req& r = pool[i];
    const uint32_t* lr = pairs.data() + r.off;
    bool short_only = r.q <= 4096;
    for (uint32_t j = 0; j < r.q && short_only; ++j) short_only = lr[2 * j + 1] - lr[2 * j] < (1u << 20);
    if (!short_only) continue;
    std::vector<uint32_t> got(r.q), want(r.q);
    uint64_t tk = 0;
    uint32_t nd = 0;
    if (rmq_server_submit(srv, ids[r.array], lr, r.q, nullptr, got.data(), &tk) != RMQ_OK ||
        rmq_server_drain(srv, done_buf.data(), static_cast<uint32_t>(done_buf.size()), &nd) != RMQ_OK)
      die("check request", RMQ_EDEVICE);
    uint32_t *dq = nullptr, *di = nullptr;
    if (hipMalloc(reinterpret_cast<void**>(&dq), r.q * 8ull) != hipSuccess ||
        hipMalloc(reinterpret_cast<void**>(&di), r.q * 4ull) != hipSuccess)
      die("hipMalloc", RMQ_ENOMEM);
    (void)hipMemcpy(dq, lr, r.q * 8ull, hipMemcpyHostToDevice);
    st = rmq_gpu_base(arrays[r.array], specs[r.array].n, dq, r.q, nullptr, di, nullptr);
    if (st != RMQ_OK || hipDeviceSynchronize() != hipSuccess) die("rmq_gpu_base", st);
    (void)hipMemcpy(want.data(), di, r.q * 4ull, hipMemcpyDeviceToHost);
    (void)hipFree(dq);
    (void)hipFree(di);
    wrong += got == want ? 0u : 1u;
    ++checked;
  }
  printf("{\"check\": \"%u requests against GPU BASE, %u differ\"}\n", checked, wrong);
  rmq_server_destroy(srv);
  for (float* p : arrays) (void)hipFree(p);
  return wrong ? 1 : 0;
}
