// tcsem_fdtd -- drop-in driver for the TCSEM_FDTD time-stepping path on one MI355X.
//
// Behaves like the reference executable (main.cu:29-310): reads FDTD_setup.in, Cond.bin,
// Receiver.in and Transmitter.in from the working directory (or --dir), runs every transmitter to
// the last receiver sample time, and writes Tx<id>.bout (int32 nr, ncomp, nt, then FP64
// [nr][ncomp][nt]) next to the inputs.  Differences, by design:
//   * the step counter, sample counters and sample-time state restart for every transmitter (the
//     reference keeps them across transmitters, main.cu:76-78,212-217, so its Tx 2.. write
//     uninitialised traces);
//   * transmitters are advanced --batch at a time in one set of launches (independent field sets
//     sharing the conductivity);
//   * --precision 32 runs FP32 fields and air-plane FFT (FP64 schedule, coefficients, receivers);
//   * --autotune times every march tiling for the batch size first and keeps the fastest;
//   * --graph K replays sample-free stretches of K steps as HIP graphs (same results).
// All device work goes through the C ABI of libtcsem.so; errors print the library message and
// exit with the (positive) error code.
#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/tcsem.h"

namespace {

int die(int code, const char *what) {
  std::fprintf(stderr, "tcsem_fdtd: %s failed (%d): %s\n", what, code, tcsem_last_error());
  return code < 0 ? -code : 1;
}

void usage() {
  std::printf(
      "usage: tcsem_fdtd [--dir D] [--precision 64|32] [--batch B] [--max-steps S] [--autotune] [--graph K]\n"
      "                  [--quiet]\n"
      "  Reads FDTD_setup.in, Cond.bin, Receiver.in, Transmitter.in from D (default: .) and writes\n"
      "  Tx<id>.bout there.  --batch: transmitters advanced together (default 8).\n"
      "  --max-steps: cap on steps per transmitter (default: until the last sample).\n"
      "  --autotune: time every march tiling for the batch size first and keep the fastest.\n"
      "  --graph K: replay sample-free stretches of K steps (even, <= 32) as HIP graphs.\n");
}

}  // namespace

int main(int argc, char **argv) {
  std::string dir = ".";
  int precision = TCSEM_FP64, batch = 8, max_steps = 0, graph = 0;
  bool quiet = false, tune = false;
  for (int a = 1; a < argc; ++a) {
    const std::string s = argv[a];
    auto next = [&](const char *name) -> const char * {
      if (a + 1 >= argc) {
        std::fprintf(stderr, "tcsem_fdtd: %s needs a value\n", name);
        std::exit(1);
      }
      return argv[++a];
    };
    if (s == "--dir") dir = next("--dir");
    else if (s == "--precision") precision = std::atoi(next("--precision"));
    else if (s == "--batch") batch = std::atoi(next("--batch"));
    else if (s == "--max-steps") max_steps = std::atoi(next("--max-steps"));
    else if (s == "--quiet") quiet = true;
    else if (s == "--autotune") tune = true;
    else if (s == "--graph") graph = std::atoi(next("--graph"));
    else if (s == "-h" || s == "--help") {
      usage();
      return 0;
    } else {
      usage();
      return 1;
    }
  }
  if ((precision != TCSEM_FP64 && precision != TCSEM_FP32) || batch < 1) {
    usage();
    return 1;
  }
  auto path = [&](const char *f) { return dir + "/" + f; };

  tcsem_setup setup;
  int rc = tcsem_read_setup(path("FDTD_setup.in").c_str(), &setup);
  if (rc) return die(rc, "reading FDTD_setup.in");
  tcsem_grid_info gi;
  if ((rc = tcsem_grid_info_of(&setup, &gi))) return die(rc, "grid setup");
  const size_t ncell = (size_t)gi.L * gi.M * gi.N;
  std::vector<double> cond(ncell);
  if ((rc = tcsem_read_cond(path("Cond.bin").c_str(), ncell, cond.data()))) return die(rc, "reading Cond.bin");

  tcsem_rxtime rt;
  int nrx = 0;
  if ((rc = tcsem_read_receivers(path("Receiver.in").c_str(), &rt, &nrx, nullptr, 0))) return die(rc, "reading Receiver.in");
  std::vector<double> rxyz(3 * (size_t)nrx);
  if ((rc = tcsem_read_receivers(path("Receiver.in").c_str(), &rt, &nrx, rxyz.data(), nrx)))
    return die(rc, "reading Receiver.in");

  int ntx = 0, nids = 0;
  if ((rc = tcsem_read_transmitters(path("Transmitter.in").c_str(), nullptr, 0, &ntx, nullptr, 0, &nids)))
    return die(rc, "reading Transmitter.in");
  std::vector<tcsem_tx> txs(ntx);
  std::vector<int> ids(nids > 0 ? nids : 1);
  if ((rc = tcsem_read_transmitters(path("Transmitter.in").c_str(), txs.data(), ntx, &ntx, ids.data(), nids, &nids)))
    return die(rc, "reading Transmitter.in");
  int ncomp = 0;
  for (int c = 0; c < 6; ++c) ncomp += rt.is_comp[c] ? 1 : 0;

  if (!quiet)
    std::printf("%s\nmodel %d x %d x %d cells, air plane %d x %d, %d receivers, %d transmitters, FP%d\n",
                tcsem_version(), gi.L, gi.M, gi.N, gi.L1, gi.M1, nrx, ntx, precision);

  tcsem_ctx *ctx = nullptr;
  if ((rc = tcsem_create(&setup, cond.data(), precision, &ctx))) return die(rc, "device setup");
  cond.clear();
  cond.shrink_to_fit();
  if ((rc = tcsem_set_receivers(ctx, &rt, nrx, rxyz.data()))) {
    tcsem_destroy(ctx);
    return die(rc, "receiver setup");
  }
  if ((rc = tcsem_set_transmitters(ctx, txs.data(), ntx))) {
    tcsem_destroy(ctx);
    return die(rc, "transmitter setup");
  }
  if (tune && ntx > 0) {
    int rows = 0, pf = 0, zc = 0;
    double ms = 0.0;
    if ((rc = tcsem_autotune(ctx, std::min(batch, ntx), 20, &rows, &pf, &zc, nullptr, &ms))) {
      tcsem_destroy(ctx);
      return die(rc, "autotune");
    }
    if (!quiet)
      std::printf("autotuned march: %d rows per tile, prefetch %d, %d z-chunk(s): %.4f ms per step\n", rows, pf, zc,
                  ms);
  }
  if (graph && (rc = tcsem_set_graph(ctx, graph))) {
    tcsem_destroy(ctx);
    return die(rc, "graph setup");
  }
  for (int first = 0; first < ntx; first += batch) {
    const int nb = std::min(batch, ntx - first);
    size_t total = 0;
    for (int b = 0; b < nb; ++b) total += (size_t)txs[first + b].nr * ncomp * rt.Nt;
    std::vector<double> traces(total > 0 ? total : 1);
    int steps = 0;
    if ((rc = tcsem_run(ctx, first, nb, max_steps, traces.data(), &steps))) {
      tcsem_destroy(ctx);
      return die(rc, "time stepping");
    }
    size_t off = 0;
    for (int b = 0; b < nb; ++b) {
      const tcsem_tx &t = txs[first + b];
      char name[64];
      std::snprintf(name, sizeof name, "Tx%d.bout", t.id);
      if ((rc = tcsem_write_bout(path(name).c_str(), t.nr, ncomp, rt.Nt, traces.data() + off))) {
        tcsem_destroy(ctx);
        return die(rc, "writing traces");
      }
      off += (size_t)t.nr * ncomp * rt.Nt;
      if (!quiet) std::printf("Tx %d: %d steps, %d receivers -> %s\n", t.id, steps, t.nr, name);
    }
  }
  tcsem_destroy(ctx);
  return 0;
}
