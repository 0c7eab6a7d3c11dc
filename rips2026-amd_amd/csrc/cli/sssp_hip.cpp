// sssp_hip: the ADDS command-line driver over libmi355sssp / libmi355graph.
//
// Same interface as the reference's driver (ADDSF/support.cpp:16-102,
// ADDSF/kernel.hip:365-393, 575-613):
//   sssp_hip [-q] [-g gpu-number] [-s start-node] [-o output-file] graph.gr
// reads a Galois .gr v1 graph, runs ADDS from the start node ADDS_RUNS times
// (default 8), prints one line per run, appends "<graph> <avg ms> <avg work>"
// to file descriptor 3 (stderr warning when fd 3 is closed), and writes
// "<node> <dist %.9g>" or "<node> INF" lines to the output file ("-" =
// stdout).  Work is the number of vertices processed.  ADDS_WORKER_BLOCKS
// overrides the persistent workgroup count.  Errors exit with status 1.
#include <getopt.h>

#include <cerrno>
#include <cfloat>
#include <climits>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <algorithm>
#include <vector>

#include "mi355graph.h"
#include "mi355sssp.h"

namespace {

int g_start = 0;
int g_device = 0;
const char* g_input = nullptr;
const char* g_output = nullptr;

void usage(const char* prog) {
  std::fprintf(stderr, "usage: %s [-q] [-g gpu-number] [-s start-node] [-o output-file] graph-file\n", prog);
}

bool parse_nonneg_int(const char* s, int* out) {
  char* end = nullptr;
  errno = 0;
  const long v = std::strtol(s, &end, 10);
  if (errno != 0 || end == s || *end != '\0' || v < 0 || v > INT_MAX) return false;
  *out = (int)v;
  return true;
}

void parse_args(int argc, char** argv) {
  int opt;
  while ((opt = getopt(argc, argv, "g:qo:s:")) != -1) {
    switch (opt) {
      case 'q':
        break;  // accepted for compatibility
      case 'o':
        g_output = optarg;
        break;
      case 'g':
        if (!parse_nonneg_int(optarg, &g_device)) {
          std::fprintf(stderr, "Invalid GPU device '%s'\n", optarg);
          std::exit(EXIT_FAILURE);
        }
        break;
      case 's':
        if (!parse_nonneg_int(optarg, &g_start)) {
          std::fprintf(stderr, "Invalid source vertex '%s'\n", optarg);
          std::exit(EXIT_FAILURE);
        }
        break;
      default:
        usage(argv[0]);
        std::exit(EXIT_FAILURE);
    }
  }
  if (optind >= argc) {
    usage(argv[0]);
    std::exit(EXIT_FAILURE);
  }
  g_input = argv[optind];
  if (optind + 1 != argc) {
    std::fprintf(stderr, "Unexpected extra argument '%s'\n", argv[optind + 1]);
    usage(argv[0]);
    std::exit(EXIT_FAILURE);
  }
}

// Integer environment knob in [lo, hi]; unset or empty gives def.  A value
// that is not a decimal integer in range ends the program, as the reference
// driver does (ADDSF/kernel.hip:34-51).
unsigned env_unsigned(const char* name, unsigned def, unsigned lo, unsigned hi) {
  const char* v = std::getenv(name);
  if (!v || !*v) return def;
  errno = 0;
  char* end = nullptr;
  const unsigned long long x = std::strtoull(v, &end, 10);
  const bool digits = std::strspn(v, "0123456789") == std::strlen(v);
  if (errno != 0 || !digits || end == v || *end != '\0' || x < lo || x > hi) {
    std::fprintf(stderr, "Invalid %s='%s'; expected an integer in [%u, %u]\n", name, v, lo, hi);
    std::exit(EXIT_FAILURE);
  }
  return (unsigned)x;
}

// ADDS_RUNS: repetitions in [1, 1000]; an invalid value is reported and the
// default used (the reference's behaviour for this knob, ADDSF/kernel.hip:365-379).
int runs_from_env() {
  constexpr int kDefaultRuns = 8;
  const char* v = std::getenv("ADDS_RUNS");
  if (!v || !*v) return kDefaultRuns;
  char* end = nullptr;
  const long x = std::strtol(v, &end, 10);
  if (end == v || *end != '\0' || x < 1 || x > 1000) {
    std::fprintf(stderr, "warning: ignoring invalid ADDS_RUNS='%s'; using %d\n", v, kDefaultRuns);
    return kDefaultRuns;
  }
  return (int)x;
}

[[noreturn]] void die_engine(const char* what) {
  std::fprintf(stderr, "%s: %s\n", what, mi355_last_error());
  std::exit(EXIT_FAILURE);
}

void summary(const char* name, double ms, unsigned long long work) {
  FILE* fd3 = std::fopen("/dev/fd/3", "a");
  if (fd3) {
    std::fprintf(fd3, "%s %.3f %llu\n", name, ms, work);
    std::fclose(fd3);
    return;
  }
  std::fprintf(stderr, "warning: file descriptor 3 is not open; benchmark summary: %s %.3f %llu\n", name, ms,
               work);
}

}  // namespace

int main(int argc, char** argv) {
  if (argc == 1) {
    usage(argv[0]);
    return EXIT_FAILURE;
  }
  parse_args(argc, argv);

  int count = 0;
  if (mi355_device_count(&count) != MI355_OK) die_engine("device query failed");
  if (g_device >= count) {
    std::fprintf(stderr, "GPU device %d is unavailable; HIP reports %d device(s)\n", g_device, count);
    return EXIT_FAILURE;
  }
  if (mi355_set_device(g_device) != MI355_OK) die_engine("cannot select the device");

  mi355g_graph* hg = nullptr;
  if (mi355g_read_galois_gr(g_input, &hg) != 0) {
    std::fprintf(stderr, "%s\n", mi355g_last_error());
    return EXIT_FAILURE;
  }
  mi355g_view_t v;
  mi355g_view(hg, &v);
  if ((int64_t)g_start >= v.rows) {
    std::fprintf(stderr, "Source vertex %d is outside [0, %lld)\n", g_start, (long long)v.rows);
    mi355g_free(hg);
    return EXIT_FAILURE;
  }
  mi355_csr_desc d{};
  d.rows = v.rows;
  d.cols = v.cols;
  d.nnz = v.nnz;
  d.rowptr = v.rowptr;
  d.col = v.col;
  d.w = v.w;
  d.offset_bits = 64;
  d.weight_type = v.weight_type == MI355G_W_U32 ? MI355_W_U32 : MI355_W_F32;
  d.orientation = MI355_OUT;
  d.memory = MI355_HOST;
  mi355_graph* g = nullptr;
  if (mi355_graph_create(&d, 0, nullptr, &g) != MI355_OK) die_engine("graph rejected");

  mi355_sssp_options o;
  mi355_sssp_options_default(&o);
  o.algorithm = MI355_ALG_ADDS;
  o.inf_mode = MI355_INF_FLT_MAX;
  // Worker pool: ADDS_RESERVED_CUS CUs are left out of it (the reference
  // keeps them for its manager kernel, ADDSF/kernel.hip:468-480; this engine
  // has no manager, so they simply stay idle), ADDS_WORKER_BLOCKS picks the
  // number of persistent workgroups up to what the remaining CUs hold.
  int32_t default_workers = 0;
  if (mi355_adds_default_workers(&default_workers) != MI355_OK || default_workers < 4)
    die_engine("device query failed");
  const unsigned per_cu = 2, cus = (unsigned)default_workers / per_cu;
  const unsigned reserved = env_unsigned("ADDS_RESERVED_CUS", 1u, 1u, cus - 1);
  const unsigned max_workers = per_cu * (cus - reserved);
  o.worker_blocks = (int32_t)env_unsigned("ADDS_WORKER_BLOCKS", max_workers, 1u, max_workers);
  // ADDS_WL_ENTRIES: the reference's worklist size (ADDSF/kernel.hip:496-510),
  // checked against the same range; the engine sizes its bucket rings from
  // the vertex count itself, so the value does not bound them.
  const uint64_t wl_min = 32ull * 65536ull * 8ull, wl_max = 536870912ull;
  const uint64_t wl_def = std::min<uint64_t>((uint64_t)std::ceil(1.5 * (double)v.nnz) + wl_min, wl_max);
  const unsigned wl_entries = env_unsigned("ADDS_WL_ENTRIES", (unsigned)wl_def, (unsigned)wl_min, (unsigned)wl_max);
  const int runs = runs_from_env();
  std::printf("workers: %d workgroup(s) of 512 threads (8 waves), reserved CUs=%u, worklist entries=%u\n",
              o.worker_blocks, reserved, wl_entries);
  void* d_dist = nullptr;
  if (mi355_device_malloc(&d_dist, 4 * (size_t)v.rows) != MI355_OK) die_engine("allocation failed");

  double total_ms = 0.0;
  unsigned long long total_work = 0;
  for (int r = 0; r < runs; ++r) {
    mi355_sssp_stats st;
    if (mi355_sssp(g, g_start, &o, d_dist, nullptr, &st, nullptr) != MI355_OK) die_engine("ADDS failed");
    total_ms += st.device_ms;
    total_work += (unsigned long long)st.vertices_processed;
    std::printf("%s run %d/%d: %.3f ms, processed vertices=%llu\n", g_input, r + 1, runs, st.device_ms,
                (unsigned long long)st.vertices_processed);
  }
  summary(g_input, total_ms / runs, total_work / (unsigned long long)runs);

  if (g_output) {
    std::vector<float> dist((size_t)v.rows);
    std::vector<uint32_t> udist;
    const bool u32 = d.weight_type == MI355_W_U32;
    if (u32) udist.resize((size_t)v.rows);
    if (mi355_memcpy_d2h(u32 ? (void*)udist.data() : (void*)dist.data(), d_dist, 4 * (size_t)v.rows, nullptr) !=
        MI355_OK)
      die_engine("copy failed");
    FILE* f = stdout;
    if (std::strcmp(g_output, "-") != 0) {
      f = std::fopen(g_output, "w");
      if (!f) {
        std::fprintf(stderr, "Unable to open output file %s: %s\n", g_output, std::strerror(errno));
        return EXIT_FAILURE;
      }
    }
    for (int64_t i = 0; i < v.rows; ++i) {
      if (u32) {
        if (udist[(size_t)i] == 0xFFFFFFFFu) std::fprintf(f, "%lld INF\n", (long long)i);
        else std::fprintf(f, "%lld %u\n", (long long)i, udist[(size_t)i]);
      } else if (dist[(size_t)i] >= FLT_MAX) {
        std::fprintf(f, "%lld INF\n", (long long)i);
      } else {
        std::fprintf(f, "%lld %.9g\n", (long long)i, dist[(size_t)i]);
      }
    }
    if (f != stdout) std::fclose(f);
  }
  mi355_device_free(d_dist);
  mi355_graph_destroy(g);
  mi355g_free(hg);
  return EXIT_SUCCESS;
}
