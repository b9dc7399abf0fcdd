// rtxrmq.cpp — command-line driver with the reference's contract (src/main.cu:71-287,
// src/tools.h:92-259, 343-617), for the hierarchical algorithms of this library:
//
//   ./rtxrmq <n> <q> <lr> <alg> [--log_bs L] [--reps R] [--dev D] [--nt T] [--seed S]
//            [--check | --trivialCheck | --randTrivialCheck] [--save-time[=file]]
//            [--save-input-data] [--nb B] [--save-power[=file]] [--log_build_threshold[=x]]
//            [--cg_size_log G] [--cg_amount_log C] [--group_lanes K] [--pack_descents] [--long]
//
// alg: 16 (XXX / GPU-RMQ CL), 20 (hierarchical vector load / GPU-RMQ VL), 24 (XXX multi load),
//      2 (GPU BASE), 0 (CPU BASE, OpenMP over queries with --nt threads); +100 returns leftmost
//      indices instead of values (100, 102, 116, 120, 124). HRMQ (1/101, an external library)
//      and the OptiX algorithms are not part of this driver.
// --check compares against GPU BASE as the reference does (src/main.cu:194-214), but
// bit-exactly (value bit patterns, indices) instead of |a-b| < 1e-4 (src/tools.h:261-264).
// The runtime flags --cg_size_log / --cg_amount_log replace the reference's compile-time
// XXX_CG_SIZE_LOG / XXX_CG_AMOUNT_LOG (CMakeLists.txt:61-67); --group_lanes picks this
// library's lanes-per-query for algs 16/24 (0 = auto; never changes results); --pack_descents
// sets RMQ_BUILD_PACK_DESCENTS (the lane solver regroups the queries still descending onto full
// wavefronts; same results).
// --long is the reference's IS_LONG build (CMakeLists.txt:57-59, src/main.cu:21-27): 64-bit
// positions, {l, r} pairs and indices; it is implied for n > 2^31 - 1 (up to 2^38).

#include <getopt.h>
#include <hip/hip_runtime_api.h>
#include <unistd.h>

#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <ctime>
#include <fstream>
#include <sstream>
#include <string>
#include <vector>

#include "../../include/rmq.h"
#include "../../include/rmq_gen.h"
#include "rtxrmq_algs.hpp"

using rtxrmq::CmdArgs;

namespace {

constexpr int kRandTrivLog = 15;  // RAND_TRIVCHECK_LOG_NUM_DIFFQ, src/main.cu:29
constexpr int kTrivLog = 10;      // TRIVCHECK_LOG_NUM_DIFFQ, src/main.cu:30

enum {
  ARG_BS = 1, ARG_NB, ARG_REPS, ARG_DEV, ARG_NT, ARG_SEED, ARG_CHECK, ARG_TIME, ARG_POWER,
  ARG_TRIVIALCHECK, ARG_THRESHOLD, ARG_SAVE_INPUTDATA, ARG_RAND_TRIVIALCHECK,
  ARG_CG_SIZE, ARG_CG_AMOUNT, ARG_GROUP_LANES, ARG_LONG, ARG_PACK
};

const char* alg_name(int alg) {
  switch (alg) {
    case 0: return "[CPU] BASE";
    case 100: return "[CPU] BASE IDX";
    case 2: return "[GPU] BASE";
    case 16: return "[GPU] XXX";
    case 20: return "[GPU] BASIC VECTOR LOAD";
    case 24: return "[GPU] XXX multi load";
    case 102: return "[GPU] BASE IDX";
    case 116: return "[GPU] XXX IDX";
    case 120: return "[GPU] BASIC VECTOR LOAD IDX";
    case 124: return "[GPU] XXX multi load IDX";
  }
  return "";
}

bool is_hier(int alg) { return alg % 100 == 16 || alg % 100 == 20 || alg % 100 == 24; }

void print_help() {
  fprintf(stderr,
          "run as ./rtxrmq <n> <q> <lr> <alg>\n\n"
          "n   = num elements\n"
          "q   = num RMQ querys\n"
          "lr  = length of range; min 1, max n\n"
          "  >0 -> value\n"
          "  -1 -> uniform distribution (large values)\n"
          "  -2 -> lognormal distribution (medium values)\n"
          "  -3 -> lognormal distribution (small values)\n"
          "  -6 -> mixed distribution (large, medium and small values)\n"
          "alg = algorithm\n"
          "   0 -> [CPU] BASE\n"
          "   2 -> [GPU] BASE\n"
          "  16 -> [GPU] XXX\n"
          "  20 -> [GPU] BASIC VECTOR LOAD\n"
          "  24 -> [GPU] XXX multi load\n"
          "  100, 102, 116, 120, 124 -> algs 0 2 16 20 24 returning indices\n\n"
          "Options:\n"
          "   --log_bs <t>              scan threshold t of the hierarchical algorithms as log2 (default: 2^15)\n"
          "   --reps <repetitions>      RMQ repeats for the avg time (default: 10)\n"
          "   --dev <device ID>         device ID (default: 0)\n"
          "   --nb <blocks>             scan threshold as n / blocks, rounded up to a power of two\n"
          "   --nt  <thread num>        number of CPU threads (algs 0 / 100)\n"
          "   --seed <seed>             seed for PRNG\n"
          "   --check                   check correctness (bit-exact against GPU BASE)\n"
          "   --randTrivialCheck        check only 2^15 random queries\n"
          "   --trivialCheck            use 2^10 distinct queries, repeated\n"
          "   --save-time=<file>        append a CSV line\n"
          "   --save-input-data         cache arrays, queries and results in $RMQ_AUX_DIR\n"
          "   --cg_size_log <g> --cg_amount_log <c> --group_lanes <k>\n"
          "   --pack_descents           lane solver with descents packed onto full wavefronts\n"
          "   --long                    64-bit positions and queries (IS_LONG; implied for n > 2^31-1)\n");
}

uint32_t log2_floor(uint64_t x) {
  uint32_t r = 0;
  while (x >>= 1) ++r;
  return r;
}

// log2 of the smallest power of two >= x (x >= 1): log2(bit_ceil(x))
uint32_t log2_ceil(uint64_t x) {
  uint32_t r = 0;
  while ((1ull << r) < x) ++r;
  return r;
}

// get_args, src/tools.h:92-259
CmdArgs get_args(int argc, char* argv[]) {
  if (argc < 5) {
    print_help();
    exit(EXIT_FAILURE);
  }
  CmdArgs a;
  a.n = atol(argv[1]);
  a.q = static_cast<uint32_t>(atol(argv[2]));
  a.lr = atoi(argv[3]);
  a.alg = atoi(argv[4]);
  if (!a.n || !a.q || !a.lr) {
    print_help();
    exit(EXIT_FAILURE);
  }
  if (a.n < 1 || a.n > (1ll << 38)) {
    fprintf(stderr, "Error: n=%ld is a non positive and therefore invalid array size!\n", (long)a.n);
    exit(EXIT_FAILURE);
  }
  if (a.lr > a.n) {
    fprintf(stderr, "Error: lr=%i > n=%ld  (lr must be between '1' and 'n')\n", a.lr, (long)a.n);
    exit(EXIT_FAILURE);
  }
  a.log_bs = 15;
  a.nb = a.n >> a.log_bs;
  a.seed = static_cast<int>(time(nullptr));
  static struct option opts[] = {
      {"log_bs", required_argument, 0, ARG_BS},
      {"nb", required_argument, 0, ARG_NB},
      {"reps", required_argument, 0, ARG_REPS},
      {"dev", required_argument, 0, ARG_DEV},
      {"nt", required_argument, 0, ARG_NT},
      {"seed", required_argument, 0, ARG_SEED},
      {"check", no_argument, 0, ARG_CHECK},
      {"trivialCheck", no_argument, 0, ARG_TRIVIALCHECK},
      {"save-input-data", no_argument, 0, ARG_SAVE_INPUTDATA},
      {"randTrivialCheck", no_argument, 0, ARG_RAND_TRIVIALCHECK},
      {"save-time", optional_argument, 0, ARG_TIME},
      {"save-power", optional_argument, 0, ARG_POWER},
      {"log_build_threshold", optional_argument, 0, ARG_THRESHOLD},
      {"cg_size_log", required_argument, 0, ARG_CG_SIZE},
      {"cg_amount_log", required_argument, 0, ARG_CG_AMOUNT},
      {"group_lanes", required_argument, 0, ARG_GROUP_LANES},
      {"pack_descents", no_argument, 0, ARG_PACK},
      {"long", no_argument, 0, ARG_LONG},
      {0, 0, 0, 0}};
  int opt, idx;
  while ((opt = getopt_long(argc, argv, "123456", opts, &idx)) != -1) {
    if (isdigit(opt)) continue;
    switch (opt) {
      case ARG_BS: {
        const int v = atoi(optarg);
        // clamp to log2(bit_floor(n)) when 2^v > n, src/tools.h:160
        a.log_bs = (a.n < (1ll << v)) ? log2_floor(static_cast<uint64_t>(a.n)) : static_cast<uint32_t>(v);
        a.nb = a.n >> a.log_bs;
        break;
      }
      case ARG_THRESHOLD:
        if (optarg) a.log_build_threshold = atoi(optarg);  // parsed but unused, as in the reference
        break;
      case ARG_NB: {  // src/tools.h:168-175: t = bit_ceil(n / nb)
        const long v = atol(optarg);
        if (v <= 0) {
          fprintf(stderr, "Error: --nb %s must be a positive number of blocks\n", optarg);
          exit(EXIT_FAILURE);
        }
        a.nb = std::min<int64_t>(a.n, v);
        a.log_bs = log2_ceil(static_cast<uint64_t>(a.n / a.nb));
        break;
      }
      case ARG_REPS: a.reps = atoi(optarg); break;
      case ARG_DEV: a.dev = atoi(optarg); break;
      case ARG_NT: a.nt = atoi(optarg); break;
      case ARG_SEED: a.seed = atoi(optarg); break;
      case ARG_CHECK: a.check = 1; a.trivialCheck = 0; a.randTrivialCheck = 0; break;
      case ARG_TRIVIALCHECK: a.trivialCheck = 1; a.check = 0; a.randTrivialCheck = 0; break;
      case ARG_RAND_TRIVIALCHECK: a.randTrivialCheck = 1; a.check = 0; a.trivialCheck = 0; break;
      case ARG_SAVE_INPUTDATA: a.save_input_data = 1; break;
      case ARG_TIME: a.save_time = 1; if (optarg) a.time_file = optarg; break;
      case ARG_POWER: a.save_power = 1; if (optarg) a.power_file = optarg; break;
      case ARG_CG_SIZE: a.cg_size_log = static_cast<uint32_t>(atoi(optarg)); break;
      case ARG_CG_AMOUNT: a.cg_amount_log = static_cast<uint32_t>(atoi(optarg)); break;
      case ARG_GROUP_LANES: a.group_lanes = static_cast<uint32_t>(atoi(optarg)); break;
      case ARG_PACK: a.pack_descents = 1; break;
      case ARG_LONG: a.is_long = 1; break;
      default: break;
    }
  }
  if (a.n > 0x7FFFFFFFll) a.is_long = 1;
  if (!is_hier(a.alg)) {  // src/tools.h:221-235
    a.log_bs = 0;
    a.nb = 0;
  }
  if (!alg_name(a.alg)[0]) {
    fprintf(stderr, "Error: alg %d is not provided by this driver (hierarchical algs 16/20/24, "
                    "GPU BASE 2, CPU BASE 0, and their +100 index forms)\n", a.alg);
    exit(EXIT_FAILURE);
  }
  printf("Params:\n   reps = %i\n   seed = %i\n   dev  = %i\n   n    = %ld (~%f GB, float)\n"
         "   bs   = %ld\n   nb   = %ld\n   q    = %u (~%f GB, int2)\n   lr   = %i\n"
         "   nt   = %i CPU threads\n   alg  = %i (%s)\n\n",
         a.reps, a.seed, a.dev, (long)a.n, 4.0 * a.n / 1e9, (long)(1ll << a.log_bs), (long)a.nb, a.q,
         8.0 * a.q / 1e9, a.lr, a.nt, a.alg, alg_name(a.alg));
  return a;
}

// ---- --save-input-data cache (src/tools.h:526-617), directory from $RMQ_AUX_DIR -------------
std::string aux_dir() {
  const char* d = getenv("RMQ_AUX_DIR");
  std::string s = d ? d : "./aux_data/";
  if (!s.empty() && s.back() != '/') s += '/';
  return s;
}

std::string make_filename(const std::string& prefix, int64_t n, uint32_t q, int lr, int seed,
                          const CmdArgs& a) {
  std::stringstream s;
  s << prefix << "_n-" << n << "_q-" << q << "_lr-" << lr << "_seed-" << seed << "_trivCheck-"
    << a.trivialCheck << "_randTrivCheck-" << a.randTrivialCheck << ".bin";
  return s.str();
}

bool file_exists(const std::string& f) {
  std::ifstream in(f, std::ios::binary);
  return in.good();
}

bool load_file(const std::string& f, void* dst, size_t bytes) {
  std::ifstream in(f, std::ios::binary);
  if (!in) return false;
  printf("\nReading data in file %s\n", f.c_str());
  in.read(static_cast<char*>(dst), static_cast<std::streamsize>(bytes));
  return static_cast<size_t>(in.gcount()) == bytes;
}

void save_file(const std::string& f, const void* src, size_t bytes) {
  std::ofstream out(f, std::ios::binary);
  if (!out) {
    fprintf(stderr, "Cannot open file for writing: %s\n", f.c_str());
    return;
  }
  printf("\nWriting data in file %s\n", f.c_str());
  out.write(static_cast<const char*>(src), static_cast<std::streamsize>(bytes));
}

std::vector<float> get_or_generate_array(const CmdArgs& a) {
  std::vector<float> A(static_cast<size_t>(a.n));
  std::stringstream s;
  s << aux_dir() << "array_n-" << a.n << "_seed-" << a.seed << ".bin";
  if (load_file(s.str(), A.data(), A.size() * 4)) return A;
  rmq_gen_array(static_cast<uint64_t>(a.n), a.seed, A.data());
  if (a.save_input_data) save_file(s.str(), A.data(), A.size() * 4);
  return A;
}

int gen_queries_p(const CmdArgs& a, uint32_t* out) {
  return rmq_gen_queries(a.q, a.lr, static_cast<uint64_t>(a.n), a.seed, a.trivialCheck, out);
}
int gen_queries_p(const CmdArgs& a, uint64_t* out) {
  return rmq_gen_queries64(a.q, a.lr, static_cast<uint64_t>(a.n), a.seed, a.trivialCheck, out);
}

// {l, r} pairs of the position type P (the reference's n2Type: uint2 or ulong2).
template <typename P>
std::vector<P> get_or_generate_queries(const CmdArgs& a) {
  std::vector<P> Q(2ull * a.q);
  const std::string f = make_filename(aux_dir() + "queries", a.n, a.q, a.lr, a.seed, a);
  if (load_file(f, Q.data(), Q.size() * sizeof(P))) return Q;
  if (gen_queries_p(a, Q.data()) != 0) {
    print_help();
    exit(EXIT_FAILURE);
  }
  if (a.save_input_data) save_file(f, Q.data(), Q.size() * sizeof(P));
  return Q;
}

// ---- CSV (src/tools.h:343-423, 488-524) -------------------------------------------------------
void create_header(const CmdArgs& a, bool xxx) {
  if (!a.save_time) return;
  const char* fn = rtxrmq::csv_path(a);
  if (FILE* probe = fopen(fn, "r")) {
    fclose(probe);
    return;
  }
  if (FILE* fp = fopen(fn, "a")) {
    if (xxx)
      fprintf(fp, "dev,alg,nt,reps,n,bs,nb,q,lr,GPU_BSIZE,CG_GROUP_SIZE,CG_MEM_ALIGNMENT,timestamp,"
                  "scan_threshold,XXX_CG_SIZE_LOG,XXX_CG_AMOUNT_LOG,t,q/s,ns/q,construction,"
                  "outbuffer,tempbuffer,freeGPUMemCorrect,freeGPUMem,checkResult\n");
    else
      fprintf(fp, "dev,alg,nt,reps,n,bs,nb,q,lr,GPU_BSIZE,CG_GROUP_SIZE,CG_MEM_ALIGNMENT,timestamp,"
                  "t,q/s,ns/q,construction,outbuffer,tempbuffer,freeGPUMemCorrect,freeGPUMem,"
                  "checkResult\n");
    fclose(fp);
  }
}

void write_params(const CmdArgs& a) {
  if (!a.save_time) return;
  hipDeviceProp_t prop;
  char device[256] = "unknown";
  const bool cpu = a.alg % 100 == 0;
  if (cpu) {  // the CPU algorithms name the host (src/tools.h:359-365)
    char host[64] = "";
    (void)gethostname(host, sizeof host - 1);
    snprintf(device, sizeof device, "CPU %s", host);
  } else if (hipGetDeviceProperties(&prop, a.dev) == hipSuccess) {
    snprintf(device, sizeof device, "%s", prop.name);
  }
  char ts[32];
  std::time_t now = std::time(nullptr);
  std::strftime(ts, sizeof ts, "%Y-%m-%dT%H:%M:%SZ", std::localtime(&now));
  FILE* fp = fopen(rtxrmq::csv_path(a), "a");
  if (!fp) return;
  const int block = cpu ? 0 : 256;  // workgroup size of this library's kernels (the reference's BSIZE)
  if (is_hier(a.alg)) {
    fprintf(fp, "%s,%s,%i,%i,%ld,%i,%ld,%u,%d,%i,%i,%i,%s,%lld,%u,%u,", device, alg_name(a.alg),
            a.nt, a.reps, (long)a.n, 1, 0L, a.q, a.lr, block, 0, 0, ts, 1ll << a.log_bs,
            a.alg % 100 != 20 ? a.cg_size_log : 0u, a.cg_amount_log);
  } else {
    fprintf(fp, "%s,%s,%i,%i,%ld,%lld,%ld,%u,%d,%i,%i,%i,%s,", device, alg_name(a.alg), a.nt,
            a.reps, (long)a.n, 1ll << a.log_bs, (long)a.nb, a.q, a.lr, block, 0, 0, ts);
  }
  fclose(fp);
}

void write_check_result(int check_result, const CmdArgs& a) {
  if (!a.save_time) return;
  if (a.trivialCheck) check_result += 2;
  else if (a.randTrivialCheck) check_result += 4;
  if (FILE* fp = fopen(rtxrmq::csv_path(a), "a")) {
    fprintf(fp, ",%i\n", check_result);
    fclose(fp);
  }
}

// Expected answers from GPU BASE over `qs` queries (value bits and leftmost indices).
template <typename P>
void expected(const float* dA, int64_t n, const std::vector<P>& qs, std::vector<float>& val,
              std::vector<P>& idx) {
  const uint64_t q = qs.size() / 2;
  val.assign(q, 0.f);
  idx.assign(q, 0u);
  void *dq = nullptr, *dv = nullptr, *di = nullptr;
  if (hipMalloc(&dq, qs.size() * sizeof(P) + 16) != hipSuccess || hipMalloc(&dv, q * 4 + 4) != hipSuccess ||
      hipMalloc(&di, q * sizeof(P) + 8) != hipSuccess) {
    fprintf(stderr, "Error: out of device memory for --check\n");
    exit(EXIT_FAILURE);
  }
  (void)hipMemcpy(dq, qs.data(), qs.size() * sizeof(P), hipMemcpyHostToDevice);
  const rmq_status st = rtxrmq::base_p(dA, n, static_cast<const P*>(dq), static_cast<uint32_t>(q),
                                       static_cast<float*>(dv), static_cast<P*>(di));
  if (st != RMQ_OK || hipDeviceSynchronize() != hipSuccess) {
    fprintf(stderr, "Error: GPU BASE failed: %s\n", rmq_status_str(st));
    exit(EXIT_FAILURE);
  }
  (void)hipMemcpy(val.data(), dv, q * 4, hipMemcpyDeviceToHost);
  (void)hipMemcpy(idx.data(), di, q * sizeof(P), hipMemcpyDeviceToHost);
  (void)hipFree(dq);
  (void)hipFree(dv);
  (void)hipFree(di);
}

// Bit-exact comparison of result[pos[i]] with expected[i] (values by bit pattern).
template <typename P>
bool compare(const std::vector<float>& A, const std::vector<P>& Q, const void* result,
             bool idx_mode, const std::vector<float>& ev, const std::vector<P>& ei,
             const std::vector<uint32_t>& pos) {
  uint64_t wrong = 0;
  for (size_t i = 0; i < pos.size(); ++i) {
    const uint32_t k = pos[i];
    bool ok;
    if (idx_mode) {
      ok = static_cast<const P*>(result)[k] == ei[i];
    } else {
      uint32_t a, b;
      memcpy(&a, static_cast<const float*>(result) + k, 4);
      memcpy(&b, &ev[i], 4);
      ok = a == b;
    }
    if (!ok) {
      if (wrong < 10) {
        printf("Error on %u-th query [%llu,%llu]: expected idx %llu (value %.10f)\n", k,
               (unsigned long long)Q[2ull * k], (unsigned long long)Q[2ull * k + 1],
               (unsigned long long)ei[i], A[ei[i]]);
      }
      ++wrong;
    }
  }
  if (wrong) printf("%.2f%% wrong.\n", 100.0 * wrong / pos.size());
  return wrong == 0;
}

template <typename P>
int run(CmdArgs a) {
  create_header(a, is_hier(a.alg));
  if (hipSetDevice(a.dev) != hipSuccess) {
    fprintf(stderr, "Error: cannot select device %d\n", a.dev);
    exit(EXIT_FAILURE);
  }

  rtxrmq::Timer timer;
  printf("Generating n=%-10ld values............", (long)a.n);
  fflush(stdout);
  std::vector<float> A = get_or_generate_array(a);
  float* dA = nullptr;
  if (hipMalloc(reinterpret_cast<void**>(&dA), A.size() * 4) != hipSuccess) {
    fprintf(stderr, "Error: out of device memory for the array\n");
    exit(EXIT_FAILURE);
  }
  (void)hipMemcpy(dA, A.data(), A.size() * 4, hipMemcpyHostToDevice);
  printf("done: %f secs\n", timer.ms() / 1000.0);
  timer.restart();
  printf("Generating q=%-10u queries...........", a.q);
  fflush(stdout);
  std::vector<P> Q = get_or_generate_queries<P>(a);
  P* dQ = nullptr;
  if (hipMalloc(reinterpret_cast<void**>(&dQ), Q.size() * sizeof(P)) != hipSuccess) {
    fprintf(stderr, "Error: out of device memory for the queries\n");
    exit(EXIT_FAILURE);
  }
  (void)hipMemcpy(dQ, Q.data(), Q.size() * sizeof(P), hipMemcpyHostToDevice);
  printf("done: %f secs\n", timer.ms() / 1000.0);
  write_params(a);

  const bool idx_mode = a.alg >= 100;
  const int base_alg = a.alg % 100;
  const uint64_t t = 1ull << a.log_bs;
  void* out = nullptr;
  switch (base_alg) {
    case 16:
      out = idx_mode ? static_cast<void*>(rtxrmq::hierarchical_rmq_idx(16, a.n, t, a.q, dA, dQ, a))
                     : rtxrmq::xxx_rmq(a.n, t, a.q, dA, dQ, a);
      break;
    case 24:
      out = idx_mode ? static_cast<void*>(rtxrmq::hierarchical_rmq_idx(24, a.n, t, a.q, dA, dQ, a))
                     : rtxrmq::xxx_rmq_multi_load(a.n, t, a.q, dA, dQ, a);
      break;
    case 20:
      out = idx_mode ? static_cast<void*>(rtxrmq::hierarchical_rmq_idx(20, a.n, t, a.q, dA, dQ, a))
                     : rtxrmq::hierarchical_vector_load(a.n, t, a.q, dA, dQ, a);
      break;
    case 2:
      out = rtxrmq::gpu_rmq_basic(a.n, a.q, dA, dQ, a, idx_mode);
      break;
    case 0:
      out = idx_mode ? static_cast<void*>(rtxrmq::cpu_rmq_idx<P>(a.q, A.data(), Q.data(), a))
                     : rtxrmq::cpu_rmq<P>(a.q, A.data(), Q.data(), a);
      break;
  }

  const int save_time = a.save_time;
  if (a.check || a.trivialCheck || a.randTrivialCheck) {
    printf(a.check ? "\nCHECKING RESULT:\n"
                   : (a.trivialCheck ? "\nTRIVIALLY CHECKING RESULT:\n"
                                     : "\nRANDOMLY TRIVIALLY CHECKING RESULT:\n"));
    // the queries whose expected answers GPU BASE computes (src/main.cu:194-283): all of them
    // (--check), the first 2^10 (--trivialCheck: the queries repeat with that period,
    // src/rand.cuh:172-179), or a sample of 2^15 (--randTrivialCheck); pos[i] is the query
    // expected answer i is compared with (every query for --trivialCheck, expected answer
    // i & (2^10 - 1), as fill_repeatedly does)
    std::vector<uint32_t> pos;
    std::vector<P> qs;
    uint64_t k = a.q;
    if (a.trivialCheck) {
      k = std::min<uint64_t>(a.q, 1ull << kTrivLog);
    } else if (a.randTrivialCheck) {
      k = std::min<uint64_t>(a.q, 1ull << kRandTrivLog);
      pos.resize(k);
      rmq_gen_sample_indices(a.q, k, a.seed, pos.data());
    }
    if (pos.empty()) {
      pos.resize(k);
      for (uint32_t i = 0; i < k; ++i) pos[i] = i;
    }
    qs.resize(k * 2);
    for (size_t i = 0; i < k; ++i) {
      qs[2 * i] = Q[2ull * pos[i]];
      qs[2 * i + 1] = Q[2ull * pos[i] + 1];
    }
    std::vector<float> ev;
    std::vector<P> ei;
    // expected answers from the $RMQ_AUX_DIR cache when present (values need both files), else
    // from GPU BASE, cached under --save-input-data; the names carry the checked count
    // (get_or_generate_expected_val / _index, src/cuda_methods.cuh:175-210)
    const std::string fv = make_filename(aux_dir() + "results_value", a.n, static_cast<uint32_t>(k), a.lr, a.seed, a);
    const std::string fi = make_filename(aux_dir() + "results_index", a.n, static_cast<uint32_t>(k), a.lr, a.seed, a);
    ev.resize(k);
    ei.resize(k);
    const bool cached = idx_mode ? load_file(fi, ei.data(), k * sizeof(P))
                                 : (file_exists(fv) && file_exists(fi) && load_file(fi, ei.data(), k * sizeof(P)) &&
                                    load_file(fv, ev.data(), k * 4));
    printf("Checking result...........................");
    fflush(stdout);
    if (!cached) {
      expected(dA, a.n, qs, ev, ei);
      if (a.save_input_data) {
        if (!idx_mode) save_file(fv, ev.data(), ev.size() * 4);
        save_file(fi, ei.data(), ei.size() * sizeof(P));
      }
    } else if (idx_mode) {
      for (size_t i = 0; i < k; ++i) ev[i] = A[ei[i]];  // for the messages only
    }
    if (a.trivialCheck) {  // every query against expected answer i & (2^10 - 1)
      const uint64_t period = k;
      std::vector<float> rv(a.q);
      std::vector<P> ri(a.q);
      pos.resize(a.q);
      for (uint32_t i = 0; i < a.q; ++i) {
        pos[i] = i;
        rv[i] = ev[i % period];
        ri[i] = ei[i % period];
      }
      ev.swap(rv);
      ei.swap(ri);
    }
    const bool pass = compare(A, Q, out, idx_mode, ev, ei, pos);
    a.save_time = save_time;
    write_check_result(pass ? 1 : 0, a);
    printf("%s\n", pass ? "pass" : "failed");
    free(out);
    (void)hipFree(dA);
    (void)hipFree(dQ);
    printf("Benchmark Finished\n");
    return pass ? 0 : 1;
  }
  write_check_result(-1, a);
  free(out);
  (void)hipFree(dA);
  (void)hipFree(dQ);
  printf("Benchmark Finished\n");
  return 0;
}

}  // namespace

int main(int argc, char* argv[]) {
  printf("----------------------------------\n");
  printf("   rtxrmq driver (MI355X, gfx950)  \n");
  printf("----------------------------------\n");
  const CmdArgs a = get_args(argc, argv);
  return a.is_long ? run<uint64_t>(a) : run<uint32_t>(a);
}
