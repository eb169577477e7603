// main.cpp -- clwilson command line (the reference's main.cpp + the result
// file handling of cl_wilson.cpp), on top of the C ABI of libclwilson.so.
//
//   clwilson -p <pmin> -P <pmax>   test primes in [pmin, pmax) -> results.txt
//   clwilson -s | --test           self test on the reference's 7 ranges
//   clwilson -r ...                also verify against goodWilsonResults.txt
//   clwilson -d <n> | --device <n> GPU index (default 0)
#include <getopt.h>
#include <inttypes.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

#include <algorithm>
#include <string>
#include <vector>

#include "boinc_stub.h"
#include "clwilson.h"
#include "putil.h"

static const uint64_t kMaxP = 0xFFFFFFFFFFFFFFFFull / 4;  // cl_wilson.h:9
static const char* kResultFile = "results.txt";
static const char* kGoodFile = "goodWilsonResults.txt";
static const uint64_t kGoodFileSum = 0x6659912234F44428ull;  // cl_wilson.cpp:970
static const char* kStateA = "stateA.ckp";  // cl_wilson.h:5-6
static const char* kStateB = "stateB.ckp";

static void help() {
  printf("clwilson: search for Wilson primes (p-1)! = -1 (mod p^2) on an AMD Instinct GPU\n");
  printf("Program usage:\n");
  printf("-p #\tStarting prime to search p\n");
  printf("-P #\tEnd prime to search P, range [-p, -P) exclusive, 5 <= -p <= p < -P <= %" PRIu64 "\n", kMaxP);
  printf("\tRequired range is <= 10e6\n");
  printf("-s\tPerform self test to verify proper operation of the program with the current GPU.\n");
  printf("-r\tVerify all results (up to 2e13) where |w_p/p| < 1/50000 with known good file %s\n", kGoodFile);
  printf("-d #\tGPU device index (default 0)\n");
  printf("-h\tPrint this help\n");
}

struct Args {
  uint64_t pmin = 0, pmax = 0;
  bool test = false, result_test = false;  // -s, -r
  int device = 0;
};

// The range's engine: a one-range batch (the batched engine's kernels,
// matrix-core run products in the 3-limb band) for ranges of at least
// kBatchMinTestPrimes test primes below 8.2e17, as clw_wilson_range chooses;
// else a session (a batch builds every sub-chunk product of a segment, which
// only pays off over many test primes; 5 limbs). Both keep one canonical
// residue per test prime in increasing p, so this program's checkpoint files
// (the reference's names and header; canonical residues and a salted sum,
// so files of the reference itself are rejected) serve either.
static const uint64_t kBatchMinTestPrimes = 65536;
struct Engine {
  clw_session* S = nullptr;
  clw_batch* B = nullptr;
  int create(clw_ctx* ctx, uint64_t pmin, uint64_t pmax, uint64_t ntp, const clw_options* opt) {
    if (ntp >= kBatchMinTestPrimes && clw_limbs_for(pmax - 1) <= 4)
      return clw_batch_create(ctx, &pmin, &pmax, 1, opt, &B);
    return clw_session_create(ctx, pmin, pmax, opt, &S);
  }
  int info(clw_range_stats* st) { return B ? clw_batch_info(B, 0, st) : clw_session_info(S, st); }
  int run(uint64_t qlo, uint64_t qhi) { return B ? clw_batch_run(B, qlo, qhi) : clw_session_run(S, qlo, qhi); }
  int residues(uint64_t* lo, uint64_t* hi, uint64_t n) {
    return B ? clw_batch_residues(B, 0, lo, hi, n) : clw_session_residues(S, lo, hi, n);
  }
  uint64_t applied() {
    if (B) return clw_batch_applied(B);
    clw_range_stats cur;
    clw_session_info(S, &cur);
    return cur.totalcount;
  }
  int resume(const uint64_t* lo, const uint64_t* hi, uint64_t n, uint64_t currp, uint64_t applied) {
    return B ? clw_batch_set_residues(B, 0, lo, hi, n, currp, applied) : clw_session_set_residues(S, lo, hi, applied);
  }
  int finish(clw_range_stats* st, uint64_t* p, uint64_t* q, uint64_t* r, uint64_t n) {
    if (!B) {
      int rc = clw_session_iterate(S);
      if (rc == CLW_OK) rc = clw_session_finalize(S, st, p, q, r, n);
      return rc;
    }
    int rc = clw_batch_iterate(B);
    if (rc == CLW_OK) rc = clw_batch_finalize(B);
    if (rc == CLW_OK) rc = clw_batch_results(B, 0, st, p, q, r, n);
    if (rc == CLW_OK) st->totalcount = clw_batch_applied(B);  // one range: every prime applied to it
    return rc;
  }
  ~Engine() {
    if (B) clw_batch_destroy(B);
    if (S) clw_session_destroy(S);
  }
};

// The -r verification (cl_wilson.cpp:902-1014, 1073-1087), then
// results.txt: one line per (near-)Wilson prime, then the checksum
// (cl_wilson.cpp:657-688, 1093-1143). 0 when both succeeded.
static int finish_outputs(const Args& a, const std::vector<uint64_t>& p, const std::vector<uint64_t>& q, uint64_t n,
                          const clw_range_stats& st, bool quiet) {
  if (a.result_test) {
    uint32_t matched = 0, listed = 0;
    if (clw_verify_results(kGoodFile, kGoodFileSum, a.pmin, a.pmax, p.data(), q.data(), n, &matched, &listed) !=
        CLW_OK) {
      fprintf(stderr, "error: %s\n", clw_last_error());
      printf("error: %s\n", clw_last_error());
      return 1;
    }
    printf("Read %u good results |w_p/p| < 1/50000 for search range from %s\n", listed, kGoodFile);
    printf("All results matched result file!\n");
  }
  char rpath[512];
  boinc_resolve_filename(kResultFile, rpath, sizeof(rpath));
  FILE* out = boinc_fopen(rpath, "w");
  if (!out) {
    printf("Cannot open %s !!!\n", kResultFile);
    return 1;
  }
  char line[128];
  uint32_t lines = 0;
  bool ok = true;
  for (uint64_t k = 0; k < n; ++k) {
    if (clw_format_result(p[k], q[k], line, sizeof(line))) {
      ok = ok && fputs(line, out) >= 0;
      ++lines;
      if (!a.result_test && !quiet) fputs(line, stdout);
    }
  }
  if (lines == 0)
    ok = ok && fprintf(out, "no results\n%016" PRIX64 "\n", st.checksum) >= 0;
  else
    ok = ok && fprintf(out, "%016" PRIX64 "\n", st.checksum) >= 0;
  if (fclose(out) != 0) ok = false;
  if (!ok) {
    printf("Cannot write to %s !!!\n", kResultFile);
    return 1;
  }
  return 0;
}

static int run_range(clw_ctx* ctx, const Args& a, clw_range_stats* st, bool quiet) {
  // setupSearch (cl_wilson.cpp:1146-1182)
  if (a.pmin == 0 || a.pmax == 0) {
    printf("-p and -P arguments are required\nuse -h for help\n");
    return 2;
  }
  if (a.pmin > a.pmax) {
    printf("pmin <= pmax is required\nuse -h for help\n");
    return 2;
  }
  if (a.pmax > a.pmin + 10000000) {
    printf("range <= 10000000 is required\nuse -h for help\n");
    return 2;
  }
  if (a.result_test && a.pmax > 20000000000000ull) {
    printf("pmax <= 2e13 is required when verifying results.\nuse -h for help\n");
    return 2;
  }
  if (!quiet) printf("Starting search at p: %" PRIu64 "\nStopping search at P: %" PRIu64 "\n", a.pmin, a.pmax);
  const uint64_t cap = 10000000;
  std::vector<uint64_t> p(cap / 10), q, r;
  // The largest range has < 700k primes; size from the host count.
  uint32_t tpcnt[3];
  uint64_t targets[3];
  if (clw_range_types(a.pmin, a.pmax, tpcnt, targets) != CLW_OK) {
    printf("error: %s\n", clw_last_error());
    return 1;
  }
  const uint64_t n = (uint64_t)tpcnt[0] + tpcnt[1] + tpcnt[2];
  if (!n) {
    printf("there are no primes to test in this range!\n");
    return 1;
  }
  if (!quiet) {
    printf("Testing %" PRIu64 " primes.  There are %u type 0 (1 mod 3) primes, %u type 1 (5 mod 12) primes, and %u "
           "type 2 (11 mod 12) primes\n",
           n, tpcnt[0], tpcnt[1], tpcnt[2]);
    printf("Factorial targets are %" PRIu64 ", %" PRIu64 ", %" PRIu64 "\n", targets[0], targets[1], targets[2]);
  }
  p.resize(n);
  q.resize(n);
  r.resize(n);
  clw_options opt{};
  opt.verbose = quiet ? 0u : 1u;
  Engine E;
  if (E.create(ctx, a.pmin, a.pmax, n, &opt) != CLW_OK) {
    printf("error: %s\n", clw_last_error());
    return 1;
  }
  clw_range_stats info;
  E.info(&info);
  // read_state (cl_wilson.cpp:1736-1777): resume from the newest valid file.
  uint64_t currp = 2;
  bool write_a_next = true;
  time_t trickle = time(nullptr);
  if (!a.test) {
    std::vector<uint64_t> lo(n), hi(n);
    clw_work_status ws;
    int which = -1, res = 0;
    char pa[512], pb[512];
    boinc_resolve_filename(kStateA, pa, sizeof(pa));
    boinc_resolve_filename(kStateB, pb, sizeof(pb));
    if (clw_ckpt_read(pa, pb, a.pmin, a.pmax, (uint32_t)n, &ws, lo.data(), hi.data(), &which, &res) !=
        CLW_OK) {
      printf("error: %s\n", clw_last_error());
      return 1;
    }
    if (res == 2) {
      printf("Workunit complete.\n");
      return 0;
    }
    if (res == 1) {
      if (E.resume(lo.data(), hi.data(), n, ws.currp, ws.totalcount) != CLW_OK) {
        printf("error: %s\n", clw_last_error());
        return 1;
      }
      currp = ws.currp;
      trickle = (time_t)ws.trickle;
      write_a_next = which == 1;
      printf("Resuming search from checkpoint. Current P: %" PRIu64 "\n", currp);
    }
  }
  auto checkpoint = [&](uint64_t cp, uint32_t done) -> int {
    std::vector<uint64_t> lo(n), hi(n);
    if (E.residues(lo.data(), hi.data(), n) != CLW_OK) return 1;
    clw_work_status ws = {a.pmin, a.pmax, cp, (uint64_t)trickle, 0, E.applied(), (uint32_t)n, done};
    char path[512];
    boinc_resolve_filename(write_a_next ? kStateA : kStateB, path, sizeof(path));
    if (clw_ckpt_write(path, &ws, lo.data(), hi.data()) == CLW_OK)
      write_a_next = !write_a_next;
    else
      fprintf(stderr, "Cannot write checkpoint to file. Continuing...\n");
    boinc_checkpoint_completed();
    return 0;
  };
  // Segment loop in windows; progress to the client and a checkpoint when it
  // asks for one (cl_wilson.cpp:1856-1929, 290-320).
  const uint64_t window = 8ull * 268492800ull;
  int rc = CLW_OK;
  for (uint64_t qlo = currp; rc == CLW_OK && qlo <= info.maxtarget;) {
    const uint64_t qhi = std::min<uint64_t>(qlo + window, info.maxtarget + 1);
    boinc_begin_critical_section();
    rc = E.run(qlo, qhi);
    boinc_end_critical_section();
    qlo = qhi;
    boinc_fraction_done((double)qlo / (double)(info.maxtarget + 1));
    if (rc == CLW_OK && !a.test && boinc_time_to_checkpoint()) {
      checkpoint(qlo, 0);
      if (!quiet) printf("Checkpoint, Current P: %" PRIu64 ", %.1f%% done\n", qlo, 100.0 * boinc_get_fraction_done());
    }
  }
  if (rc == CLW_OK) rc = E.finish(st, p.data(), q.data(), r.data(), n);
  if (rc != CLW_OK) {
    fprintf(stderr, "error: %s\n", clw_last_error());
    printf("error: %s\n", clw_last_error());
    return 1;
  }
  // As getResults + finalizeResults + the done checkpoint of
  // cl_wilson.cpp:1949-1956, in one critical section: the -r check, then
  // results.txt (every write checked), and only once it is closed the
  // checkpoint that marks the workunit done, so a failure or a kill before
  // that point leaves a workunit that is searched again, never one that
  // reports complete without its results.
  boinc_begin_critical_section();
  const int out_rc = finish_outputs(a, p, q, n, *st, quiet);
  if (out_rc == 0 && !a.test) checkpoint(info.maxtarget + 1, 1);
  boinc_end_critical_section();
  if (out_rc != 0) return out_rc;
  if (!quiet) {
    printf("Search finished in %.2f sec.\n", st->seconds);
    printf("results %u, total power table primes generated %" PRIu64 ", checksum %016" PRIX64 "\n", st->resultcount,
           st->totalcount, st->checksum);
  }
  return 0;
}

struct Kat {
  uint64_t pmin, pmax;
  uint32_t results;
  uint64_t checksum, pi_maxt, lastp;
  int64_t lastw;
};

// cl_wilson.cpp:1998-2139 run_test: the reference's seven known answers.
static const Kat kKats[7] = {
    {1239053554603ull, 1239053554604ull, 1, 0x00000240FAB1A752ull, 8257082014ull, 1239053554603ull, -4},
    {1108967825921ull, 1108967825922ull, 1, 0x0000010233A2220Dull, 10956003002ull, 1108967825921ull, 12},
    {5609877309359ull, 5609877309360ull, 1, 0x00000A344D7D0F58ull, 101542897873ull, 5609877309359ull, -6},
    {16556218163369ull, 16556218163370ull, 1, 0x00000F0ECB80A0ABull, 147755473426ull, 16556218163369ull, 2},
    {200, 564, 57, 0x00000000000080A3ull, 30, 563, 0},
    {86000000, 87467200, 1, 0x0000097C61AB0943ull, 2604536, 87467099, -2},
    {17524177394450ull, 17524177394618ull, 1, 0x00005B54B4CBBC47ull, 304620766446ull, 17524177394617ull, 256},
};

static int self_test(clw_ctx* ctx) {
  printf("Beginning self test of 7 ranges.\n\n");
  const time_t t0 = time(nullptr);
  int good = 0;
  for (int k = 0; k < 7; ++k) {
    Args a;
    a.pmin = kKats[k].pmin;
    a.pmax = kKats[k].pmax;
    a.test = true;
    clw_range_stats st;
    const int rc = run_range(ctx, a, &st, true);
    const bool ok = rc == 0 && st.resultcount == kKats[k].results && st.checksum == kKats[k].checksum &&
                    st.totalcount == kKats[k].pi_maxt && st.lastp == kKats[k].lastp && st.lastw == kKats[k].lastw;
    printf("test case %d %s. (%.2f s)\n\n", k + 1, ok ? "passed" : "failed", st.seconds);
    good += ok;
  }
  printf(good == 7 ? "All test cases completed successfully!\n" : "Self test FAILED!\n");
  printf("Elapsed time: %d sec.\n", (int)(time(nullptr) - t0));
  return good == 7 ? 0 : 1;
}

int main(int argc, char** argv) {
  Args a;
  static const struct option long_opts[] = {{"device", required_argument, 0, 'd'},
                                            {"test", no_argument, 0, 's'},
                                            {0, 0, 0, 0}};
  int opt, ind = -1;
  while ((opt = getopt_long(argc, argv, "p:P:srd:h", long_opts, &ind)) != -1) {
    int st = 0;
    switch (opt) {
      case 'p': st = clw_parse_uint64(&a.pmin, optarg, 5, kMaxP - 1); break;
      case 'P': st = clw_parse_uint64(&a.pmax, optarg, 6, kMaxP); break;
      case 's': a.test = true; break;
      case 'r': a.result_test = true; break;
      case 'd': a.device = atoi(optarg); break;
      case 'h':
      default: help(); return opt == 'h' ? 0 : 2;
    }
    if (st == -1) {
      printf("%s: invalid argument -%c %s\n", argv[0], opt, optarg);
      return 2;
    }
    if (st == -2) {
      printf("%s: out of range argument -%c %s\n", argv[0], opt, optarg);
      return 2;
    }
  }
  boinc_init();  // main.cpp:169-173
  clw_ctx* ctx = nullptr;
  if (clw_init(a.device, &ctx) != CLW_OK) {
    fprintf(stderr, "error: %s\n", clw_last_error());
    return boinc_finish(1);
  }
  clw_device_info info;
  clw_get_device_info(ctx, &info);
  printf("GPU Info:\n  Name: \t\t%s\n  Arch: \t\t%s\n  Compute Units: \t%d\n", info.name, info.arch,
         info.compute_units);
  int rc;
  if (a.test) {
    rc = self_test(ctx);
  } else {
    clw_range_stats st;
    rc = run_range(ctx, a, &st, false);
  }
  clw_shutdown(ctx);
  return boinc_finish(rc);  // main.cpp:328
}
