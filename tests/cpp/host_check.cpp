// Host-only code paths under AddressSanitizer and UndefinedBehaviorSanitizer.
//
// Built by tests/cpp/Makefile with g++ -fsanitize=address,undefined from the engine's
// host-only sources (cf_preset.cpp, cf_host.cpp, with cf_json.h and cf_tune_cache.h) and the
// CPU oracle (oracle/oracle.cpp); no GPU code.  The CPU tests run it on the reference's
// preset files, on malformed preset and tuning-cache files, and on random oracle inputs.
// A status the code returns (CF_ERR_ARG for a bad file, ...) is a normal outcome and is
// printed as "rc=<status>"; only a sanitizer report or a crash makes the exit status
// non-zero.
//
//   host_check preset <in.json> [<out.json>]  read; with <out>, write back with <in> as template
//   host_check cache-lookup <file> <key>
//   host_check cache-store <file> <key>
//   host_check oracle <n> <seed>              step, edges, triangles, pairs, histogram

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <random>
#include <vector>

#include "../../include/cellflow.h"

extern "C" {
struct orc_params {
  float radius, dt, friction, repulsion, attraction, k, balance;
  float L[3];
  int T;
  float ratio, force_multiplier;
  int max_expected_neighbors;
};
}
// the oracle's parameter block (oracle/oracle.cpp), 14 four-byte fields
static_assert(sizeof(orc_params) == 56, "orc_params layout");
extern "C" {
int orc_step(int n, const double* pos, const double* vel, const uint32_t* type,
             const int32_t* prev, const orc_params* p, const float* table, const float* rbt,
             double* pos_out, double* vel_out, double* acc_out, int32_t* count_out,
             double* fabs_out, int mode, int nthreads);
int orc_edges(int n, const float* pos, const uint32_t* type, float dist, int max_conn, int mode,
              int32_t* out, long long cap, long long* count, int nthreads);
int orc_triangles(int n, const float* pos, const uint32_t* type, float dist, int max_conn,
                  int mode, int32_t* out, long long cap, long long* count, int nthreads);
int orc_pairs(int n, const float* pos, const uint32_t* type, const orc_params* pp,
              const float* rbt, int64_t* offs, int32_t* nbr, int64_t cap, int nthreads);
int orc_hist(int n, const double* pos, const float* L, int bins, double r_max, double* hist,
             int nthreads);
}

static int preset(int argc, char** argv) {
  cf_preset p;
  cf_preset_init(&p);
  int rc = cf_preset_read(argv[2], &p);
  std::printf("read rc=%d %s\n", rc, rc ? cf_last_error_string() : "");
  if (rc == CF_OK) {
    cf_params prm;
    rc = cf_preset_params(&p, &prm);
    std::printf("params rc=%d\n", rc);
  }
  if (argc > 3) {
    rc = cf_preset_write(argv[3], &p, argv[2]);
    std::printf("write rc=%d %s\n", rc, rc ? cf_last_error_string() : "");
  }
  return 0;
}

static int cache(int argc, char** argv, bool store) {
  if (argc < 4) return 2;
  cf_tuning t{1, 3, 8, 256, 32, 1};
  double us = 0.0;
  int found = 0;
  const int rc = store ? cf_tuning_cache_store(argv[2], argv[3], &t, 46.25)
                       : cf_tuning_cache_lookup(argv[2], argv[3], &t, &us, &found);
  std::printf("rc=%d found=%d %s\n", rc, found, rc ? cf_last_error_string() : "");
  return 0;
}

static int oracle(int n, unsigned seed) {
  std::mt19937 g(seed);
  std::uniform_real_distribution<double> u(0.0, 0.3);
  const int T = 3;
  std::vector<double> pos(3 * n), vel(3 * n, 0.0), po(3 * n), vo(3 * n), ao(3 * n), fab(n);
  std::vector<uint32_t> type(n);
  std::vector<int32_t> prev(n, 0), cnt(n);
  std::vector<float> pf(3 * n), table(T * T), rbt(T);
  for (int i = 0; i < 3 * n; ++i) pf[i] = (float)(pos[i] = u(g));
  for (int i = 0; i < n; ++i) type[i] = (uint32_t)(g() % T);
  for (auto& x : table) x = (float)(u(g) * 4.0 - 0.6);
  for (auto& x : rbt) x = (float)(u(g) * 6.0 - 0.9);
  orc_params p{0.05f, 0.01f, 0.1f, 64.83f, 3.06f, 1.0f, 0.79f, {0.3f, 0.3f, 0.3f}, T, 0.3f,
               2.33f, 400};
  for (int mode = 0; mode < 2; ++mode) {
    int rc = orc_step(n, pos.data(), vel.data(), type.data(), prev.data(), &p, table.data(),
                      rbt.data(), po.data(), vo.data(), ao.data(), cnt.data(), fab.data(), mode,
                      2);
    std::printf("step mode %d rc=%d\n", mode, rc);
    std::vector<int32_t> e(2 * (size_t)n * 8);
    long long ne = 0;
    rc = orc_edges(n, pf.data(), type.data(), 0.04f, 8, mode, e.data(), (long long)n * 8, &ne, 2);
    std::printf("edges rc=%d %lld\n", rc, ne);
    std::vector<int32_t> tr(3 * (size_t)n * 64);
    long long nt = 0;
    rc = orc_triangles(n, pf.data(), type.data(), 0.04f, 8, mode, tr.data(), (long long)n * 64,
                       &nt, 2);
    std::printf("triangles rc=%d %lld\n", rc, nt);
  }
  // a pair list too small for its content must be reported, not overrun
  std::vector<int64_t> offs(n + 1);
  std::vector<int32_t> nbr(4);
  int rc = orc_pairs(n, pf.data(), type.data(), &p, rbt.data(), offs.data(), nbr.data(), 4, 2);
  std::printf("pairs (capacity 4) rc=%d\n", rc);
  nbr.assign((size_t)n * 64, 0);
  rc = orc_pairs(n, pf.data(), type.data(), &p, rbt.data(), offs.data(), nbr.data(),
                 (int64_t)nbr.size(), 2);
  std::printf("pairs rc=%d %lld\n", rc, (long long)offs[n]);
  std::vector<double> h(16);
  rc = orc_hist(n, pos.data(), p.L, 16, 0.1, h.data(), 2);
  std::printf("hist rc=%d\n", rc);
  return 0;
}

int main(int argc, char** argv) {
  if (argc < 3) {
    std::fprintf(stderr, "usage: host_check preset|cache-lookup|cache-store|oracle ...\n");
    return 2;
  }
  if (!std::strcmp(argv[1], "preset")) return preset(argc, argv);
  if (!std::strcmp(argv[1], "cache-lookup")) return cache(argc, argv, false);
  if (!std::strcmp(argv[1], "cache-store")) return cache(argc, argv, true);
  if (!std::strcmp(argv[1], "oracle") && argc > 3)
    return oracle(std::atoi(argv[2]), (unsigned)std::atoi(argv[3]));
  return 2;
}
