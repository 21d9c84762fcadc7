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
//   * --graph K replays sample-free stretches of K steps as HIP graphs (same results);
//   * several --dir decks (models of any grid sizes) run as one model batch, the n-th
//     transmitter of every deck together, each deck writing its own Tx<id>.bout.
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
      "usage: tcsem_fdtd [--dir D]... [--precision 64|32] [--batch B] [--max-steps S] [--autotune] [--graph K]\n"
      "                  [--quiet]\n"
      "  Reads FDTD_setup.in, Cond.bin, Receiver.in, Transmitter.in from D (default: .) and writes\n"
      "  Tx<id>.bout there.  --batch: transmitters advanced together (default 8).\n"
      "  --max-steps: cap on steps per transmitter (default: until the last sample).\n"
      "  --autotune: time every march tiling for the batch size first and keep the fastest.\n"
      "  --graph K: replay sample-free stretches of K steps (even, <= 32) as HIP graphs.\n"
      "  Several --dir: the decks (models of any grid size) run as one model batch, transmitter by\n"
      "  transmitter (the n-th transmitter of every deck together); each deck gets its own files.\n");
}

// One deck directory: the reference's four input files, read through the C ABI.
struct Deck {
  std::string dir;
  tcsem_setup setup{};
  tcsem_grid_info gi{};
  std::vector<double> cond;
  tcsem_rxtime rt{};
  int nrx = 0, ntx = 0, ncomp = 0;
  std::vector<double> rxyz;
  std::vector<tcsem_tx> txs;
  std::vector<int> ids;
  tcsem_ctx *ctx = nullptr;
  std::string path(const char *f) const { return dir + "/" + f; }
};

int load_deck(Deck &d) {
  int rc = tcsem_read_setup(d.path("FDTD_setup.in").c_str(), &d.setup);
  if (rc) return die(rc, ("reading " + d.path("FDTD_setup.in")).c_str());
  if ((rc = tcsem_grid_info_of(&d.setup, &d.gi))) return die(rc, "grid setup");
  const size_t ncell = (size_t)d.gi.L * d.gi.M * d.gi.N;
  d.cond.assign(ncell, 0.0);
  if ((rc = tcsem_read_cond(d.path("Cond.bin").c_str(), ncell, d.cond.data())))
    return die(rc, ("reading " + d.path("Cond.bin")).c_str());
  if ((rc = tcsem_read_receivers(d.path("Receiver.in").c_str(), &d.rt, &d.nrx, nullptr, 0)))
    return die(rc, ("reading " + d.path("Receiver.in")).c_str());
  d.rxyz.assign(3 * (size_t)d.nrx, 0.0);
  if ((rc = tcsem_read_receivers(d.path("Receiver.in").c_str(), &d.rt, &d.nrx, d.rxyz.data(), d.nrx)))
    return die(rc, "reading Receiver.in");
  int nids = 0;
  if ((rc = tcsem_read_transmitters(d.path("Transmitter.in").c_str(), nullptr, 0, &d.ntx, nullptr, 0, &nids)))
    return die(rc, ("reading " + d.path("Transmitter.in")).c_str());
  d.txs.resize(d.ntx);
  d.ids.assign(nids > 0 ? nids : 1, 0);
  if ((rc = tcsem_read_transmitters(d.path("Transmitter.in").c_str(), d.txs.data(), d.ntx, &d.ntx, d.ids.data(), nids,
                                    &nids)))
    return die(rc, "reading Transmitter.in");
  d.ncomp = 0;
  for (int c = 0; c < 6; ++c) d.ncomp += d.rt.is_comp[c] ? 1 : 0;
  return 0;
}

int setup_context(Deck &d, int precision) {
  int rc = tcsem_create(&d.setup, d.cond.data(), precision, &d.ctx);
  if (rc) return die(rc, ("device setup of " + d.dir).c_str());
  d.cond.clear();
  d.cond.shrink_to_fit();
  if ((rc = tcsem_set_receivers(d.ctx, &d.rt, d.nrx, d.rxyz.data()))) return die(rc, "receiver setup");
  if ((rc = tcsem_set_transmitters(d.ctx, d.txs.data(), d.ntx))) return die(rc, "transmitter setup");
  return 0;
}

int write_traces(const Deck &d, const tcsem_tx &t, const double *traces, int steps, bool quiet) {
  char name[64];
  std::snprintf(name, sizeof name, "Tx%d.bout", t.id);
  const int rc = tcsem_write_bout(d.path(name).c_str(), t.nr, d.ncomp, d.rt.Nt, traces);
  if (rc) return die(rc, "writing traces");
  if (!quiet) std::printf("%s: Tx %d: %d steps, %d receivers -> %s\n", d.dir.c_str(), t.id, steps, t.nr, name);
  return 0;
}

// Several decks as one model batch: for each transmitter index, every deck that has it is begun
// with that transmitter and the batch steps them together until all have sampled.
int run_model_batch(std::vector<Deck> &decks, int max_steps, bool tune, bool quiet) {
  int most = 0;
  for (const Deck &d : decks) most = std::max(most, d.ntx);
  bool tuned = !tune;
  for (int t = 0; t < most; ++t) {
    std::vector<tcsem_ctx *> members;
    std::vector<Deck *> ds;
    for (Deck &d : decks)
      if (t < d.ntx) {
        int rc = tcsem_begin(d.ctx, t, 1);
        if (rc) return die(rc, "begin");
        members.push_back(d.ctx);
        ds.push_back(&d);
      }
    tcsem_batch *b = nullptr;
    int rc = tcsem_batch_create(members.data(), (int)members.size(), &b);
    if (rc) return die(rc, "model batch");
    if (!tuned) {
      int rows = 0, pf = 0;
      double ms = 0.0;
      if ((rc = tcsem_batch_autotune(b, 10, &rows, &pf, &ms))) {
        tcsem_batch_destroy(b);
        return die(rc, "autotune");
      }
      if (!quiet) std::printf("autotuned model batch: %d rows per tile, prefetch %d: %.4f ms per step\n", rows, pf, ms);
      tuned = true;
    }
    int steps = 0;
    if ((rc = tcsem_batch_run(b, max_steps, &steps))) {
      tcsem_batch_destroy(b);
      return die(rc, "time stepping");
    }
    tcsem_batch_destroy(b);
    for (Deck *d : ds) {
      const tcsem_tx &tx = d->txs[t];
      std::vector<double> traces(std::max<size_t>((size_t)tx.nr * d->ncomp * d->rt.Nt, 1));
      if ((rc = tcsem_get_traces(d->ctx, 0, traces.data()))) return die(rc, "reading traces");
      if ((rc = write_traces(*d, tx, traces.data(), steps, quiet))) return rc;
    }
  }
  return 0;
}

}  // namespace

int main(int argc, char **argv) {
  std::vector<std::string> dirs;
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
    if (s == "--dir") dirs.push_back(next("--dir"));
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
  if (dirs.empty()) dirs.push_back(".");
  std::vector<Deck> decks(dirs.size());
  for (size_t q = 0; q < dirs.size(); ++q) {
    decks[q].dir = dirs[q];
    if (int rc = load_deck(decks[q])) return rc;
    const Deck &d = decks[q];
    if (!quiet)
      std::printf("%s: model %d x %d x %d cells, air plane %d x %d, %d receivers, %d transmitters, FP%d\n",
                  d.dir.c_str(), d.gi.L, d.gi.M, d.gi.N, d.gi.L1, d.gi.M1, d.nrx, d.ntx, precision);
  }
  if (!quiet) std::printf("%s\n", tcsem_version());
  auto cleanup = [&]() {
    for (Deck &d : decks) tcsem_destroy(d.ctx);
  };
  for (Deck &d : decks)
    if (int rc = setup_context(d, precision)) {
      cleanup();
      return rc;
    }
  if (decks.size() > 1) {
    const int rc = run_model_batch(decks, max_steps, tune, quiet);
    cleanup();
    return rc;
  }
  Deck &d = decks[0];
  tcsem_ctx *ctx = d.ctx;
  int rc = 0;
  if (tune && d.ntx > 0) {
    int rows = 0, pf = 0, zc = 0;
    double ms = 0.0;
    if ((rc = tcsem_autotune(ctx, std::min(batch, d.ntx), 20, &rows, &pf, &zc, nullptr, &ms))) {
      cleanup();
      return die(rc, "autotune");
    }
    if (!quiet)
      std::printf("autotuned march: %d rows per tile, prefetch %d, %d z-chunk(s): %.4f ms per step\n", rows, pf, zc,
                  ms);
  }
  if (graph && (rc = tcsem_set_graph(ctx, graph))) {
    cleanup();
    return die(rc, "graph setup");
  }
  for (int first = 0; first < d.ntx; first += batch) {
    const int nb = std::min(batch, d.ntx - first);
    size_t total = 0;
    for (int b = 0; b < nb; ++b) total += (size_t)d.txs[first + b].nr * d.ncomp * d.rt.Nt;
    std::vector<double> traces(total > 0 ? total : 1);
    int steps = 0;
    if ((rc = tcsem_run(ctx, first, nb, max_steps, traces.data(), &steps))) {
      cleanup();
      return die(rc, "time stepping");
    }
    size_t off = 0;
    for (int b = 0; b < nb; ++b) {
      const tcsem_tx &t = d.txs[first + b];
      if ((rc = write_traces(d, t, traces.data() + off, steps, quiet))) {
        cleanup();
        return rc;
      }
      off += (size_t)t.nr * d.ncomp * d.rt.Nt;
    }
  }
  cleanup();
  return 0;
}
