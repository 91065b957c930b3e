// Replays a recorded sequence of Numerics calls through the C++ adapter
// (include/peanoclaw_amd_numerics.hpp) on patches that live in heap vectors,
// as Peano's patches do, and writes the final heap contents.
//
//   numerics_replay TRACE BLOB OUT
//
// TRACE holds one call per line (doubles as C99 hex floats, exact):
//   N h level virtual n x y size time dt   new patch h: a leaf is initialised
//                                           (initializePatch); a virtual one is a
//                                           cell created refined whose heap vector
//                                           came from Patch::switchToVirtual (the
//                                           adapter registers it on first use)
//   S h offset                              heap uNew h <- blob[offset..] (n*n),
//                                           b <- next (n+2)^2, uOld <- 0; upload
//   V h virtual                             virtual/leaf switch of the description
//                                           (a leaf keeps its heap vector)
//   Z h                                     a virtual patch's next grid iteration:
//                                           stepUp deletes its heap vector
//                                           (switchToNonVirtual, uIndex -1), stepDown
//                                           creates a fresh zeroed one with a new
//                                           uIndex (switchToVirtual) and the
//                                           scenario's update puts b back
//                                           (Patch.cpp:732-790); the adapter sees
//                                           the new uIndex at its next call
//   R src dst | P dst                       restrictSolution | postProcessRestriction
//   C h                                     copyUNewToUOld
//   T sx sy sox soy dox doy src dst         transferGhostlayer
//   I sx sy dox doy src dst toOld toCur useTime   interpolateSolution
//   W h dim upper                           fillBoundaryLayer
//   F h cfl                                 the CFL step of h; must equal `cfl`
//   K h time dt                             the host clock of h; must equal
//   X h maxdt                               solveTimestep(maxdt)
//   D h thr minw maxw w0 w1                 getDemandedMeshWidth with the
//                                           criterion (thr, minw, maxw); must
//                                           equal (w0, w1)
//   U h offset time dt                      switch to a leaf: heap uOld (3
//                                           (n+2)^2) and uNew (3 n^2) <- blob,
//                                           clock (time, dt); upload
//   B h offset                              heap b <- blob ((n+2)^2); upload
//   Q h                                     destroyPatch (the slot is released)
// OUT receives, per live (not released) patch in creation order: time,
// timestep size, estimated next timestep size (3 doubles), then uNew after
// addPatchToSolution.
// Exit status 0 when every F, K and D matched, 1 otherwise, 2 on errors.
#include <array>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <memory>
#include <sstream>
#include <string>
#include <vector>

#include "peanoclaw_amd_numerics.hpp"

namespace harness {

// The CellDescription fields the Numerics path reads
// (dastgen/CellDescription.def).
struct CellDescription {
  int cellDescriptionIndex = -1;
  int uIndex = -1;
  std::array<double, 2> position{}, size{}, demanded{}, lower{}, upper{};
  std::array<int, 2> subdivision{};
  int level = 1;
  bool isVirtual = false;
  double time = 0.0, dt = 0.0, est = 0.0, minNeighbor = 1e300;
};

class TimeIntervals {
 public:
  explicit TimeIntervals(CellDescription* c) : c_(c) {}
  double getCurrentTime() const { return c_->time; }
  void setCurrentTime(double t) { c_->time = t; }
  double getTimestepSize() const { return c_->dt; }
  void setTimestepSize(double dt) { c_->dt = dt; }
  double getEstimatedNextTimestepSize() const { return c_->est; }
  void setEstimatedNextTimestepSize(double e) { c_->est = e; }
  double getMinimalNeighborTimeConstraint() const { return c_->minNeighbor; }

 private:
  CellDescription* c_;
};

// One heap vector per patch in the reference layout: [uNew 3 n^2]
// [uOld 3 (n+2)^2] [parameter with ghost layer (n+2)^2].
class Accessor {
 public:
  Accessor(std::vector<double>* u, int n) : u_(u), n_(n) {}
  double* getUNewArray() const { return u_->data(); }
  double* getUOldWithGhostLayerArray(int q) const {
    return u_->data() + 3 * n_ * n_ + q * (n_ + 2) * (n_ + 2);
  }
  double* getParameterWithGhostLayerArray(int) const {
    return u_->data() + 3 * n_ * n_ + 3 * (n_ + 2) * (n_ + 2);
  }

 private:
  std::vector<double>* u_;
  int n_;
};

class HeapPatch {
 public:
  HeapPatch(int key, int uIndex, int level, bool virt, int n, double x, double y, double size,
            double t, double dt)
      : n_(n), heap_((size_t)3 * n * n + 4 * (size_t)(n + 2) * (n + 2), 0.0), ti_(&c_),
        acc_(&heap_, n) {
    c_.cellDescriptionIndex = key;
    c_.uIndex = uIndex;
    c_.level = level;
    c_.isVirtual = virt;
    c_.position = {x, y};
    c_.size = {size, size};
    c_.subdivision = {n, n};
    c_.time = t;
    c_.dt = dt;
  }
  int getCellDescriptionIndex() const { return c_.cellDescriptionIndex; }
  int getUIndex() const { return c_.uIndex; }
  // switchToNonVirtual + switchToVirtual: a fresh zeroed heap vector under a
  // new uIndex, b restored (the scenario's update)
  void replaceHeapVector(int uIndex) {
    const size_t a = (size_t)n_ * n_, g = (size_t)(n_ + 2) * (n_ + 2);
    std::fill(heap_.begin(), heap_.begin() + 3 * a + 3 * g, 0.0);
    c_.uIndex = uIndex;
  }
  std::array<double, 2> getPosition() const { return c_.position; }
  std::array<double, 2> getSize() const { return c_.size; }
  std::array<int, 2> getSubdivisionFactor() const { return c_.subdivision; }
  std::array<double, 2> getDemandedMeshWidth() const { return c_.demanded; }
  std::array<double, 2> getLowerNeighboringGhostlayerBounds() const { return c_.lower; }
  std::array<double, 2> getUpperNeighboringGhostlayerBounds() const { return c_.upper; }
  int getGhostlayerWidth() const { return 1; }
  int getLevel() const { return c_.level; }
  bool isVirtual() const { return c_.isVirtual; }
  void setVirtual(bool v) { c_.isVirtual = v; }
  void setDemandedMeshWidth(double w0, double w1) { c_.demanded = {w0, w1}; }
  TimeIntervals& getTimeIntervals() { return ti_; }
  Accessor& getAccessor() { return acc_; }
  std::vector<double>& heap() { return heap_; }
  int n() const { return n_; }

 private:
  int n_;
  CellDescription c_;
  std::vector<double> heap_;
  TimeIntervals ti_;
  Accessor acc_;
};

}  // namespace harness

int main(int argc, char** argv) {
  using harness::HeapPatch;
  if (argc != 4) {
    std::fprintf(stderr, "usage: %s TRACE BLOB OUT\n", argv[0]);
    return 2;
  }
  std::vector<double> blob;
  {
    std::ifstream b(argv[2], std::ios::binary);
    if (!b) {
      std::fprintf(stderr, "cannot read %s\n", argv[2]);
      return 2;
    }
    b.seekg(0, std::ios::end);
    blob.resize((size_t)b.tellg() / sizeof(double));
    b.seekg(0);
    b.read(reinterpret_cast<char*>(blob.data()), (std::streamsize)(blob.size() * sizeof(double)));
  }
  std::ifstream tr(argv[1]);
  if (!tr) {
    std::fprintf(stderr, "cannot read %s\n", argv[1]);
    return 2;
  }
  long calls = 0, cflChecks = 0, cflMismatch = 0, clockChecks = 0, clockMismatch = 0;
  long widthChecks = 0, widthMismatch = 0;
  int nextUIndex = 0;  // DataHeap indices handed out
  std::vector<std::unique_ptr<HeapPatch>> patches;
  try {
    peanoclaw_amd::SWENumerics<HeapPatch> num;
    auto P = [&](int h) -> HeapPatch& {
      if (h < 0 || h >= (int)patches.size() || !patches[h])
        throw peanoclaw_amd::Error(PC_E_INVALID, "trace names an unknown patch");
      return *patches[h];
    };
    auto hexd = [](const std::string& s) { return std::strtod(s.c_str(), nullptr); };
    std::string line;
    while (std::getline(tr, line)) {
      if (line.empty()) continue;
      std::istringstream in(line);
      std::string op;
      in >> op;
      ++calls;
      if (op == "N") {
        int h, level, virt, n;
        std::string x, y, size, t, dt;
        in >> h >> level >> virt >> n >> x >> y >> size >> t >> dt;
        if (h != (int)patches.size()) throw peanoclaw_amd::Error(PC_E_INVALID, "handles out of order");
        patches.emplace_back(new HeapPatch(h, nextUIndex++, level, virt != 0, n, hexd(x),
                                           hexd(y), hexd(size), hexd(t), hexd(dt)));
        if (!virt) num.initializePatch(*patches.back());
      } else if (op == "S") {
        int h;
        size_t off;
        in >> h >> off;
        HeapPatch& p = P(h);
        const int n = p.n();
        const size_t a = (size_t)n * n, g = (size_t)(n + 2) * (n + 2);
        if (off + a + g > blob.size()) throw peanoclaw_amd::Error(PC_E_INVALID, "blob too short");
        std::vector<double>& u = p.heap();
        std::fill(u.begin(), u.end(), 0.0);
        std::copy(blob.begin() + off, blob.begin() + off + a, u.begin());
        std::copy(blob.begin() + off + a, blob.begin() + off + a + g, u.begin() + 3 * a + 3 * g);
        num.uploadFromHeap(p);
      } else if (op == "V") {
        int h, v;
        in >> h >> v;
        P(h).setVirtual(v != 0);
        num.push(P(h));
      } else if (op == "Z") {
        int h;
        in >> h;
        P(h).replaceHeapVector(nextUIndex++);
      } else if (op == "R") {
        int s, d;
        in >> s >> d;
        num.restrictSolution(P(s), P(d), false);
      } else if (op == "P") {
        int d;
        in >> d;
        num.postProcessRestriction(P(d), false);
      } else if (op == "C") {
        int h;
        in >> h;
        num.copyUNewToUOld(P(h));
      } else if (op == "T") {
        std::array<int, 2> sz, so, doff;
        int s, d;
        in >> sz[0] >> sz[1] >> so[0] >> so[1] >> doff[0] >> doff[1] >> s >> d;
        num.transferGhostlayer(sz, so, doff, P(s), P(d));
      } else if (op == "I") {
        std::array<int, 2> sz, doff;
        int s, d, a, b, c;
        in >> sz[0] >> sz[1] >> doff[0] >> doff[1] >> s >> d >> a >> b >> c;
        num.interpolateSolution(sz, doff, P(s), P(d), a != 0, b != 0, c != 0);
      } else if (op == "W") {
        int h, dim, up;
        in >> h >> dim >> up;
        num.fillBoundaryLayer(P(h), dim, up != 0);
      } else if (op == "F") {
        int h;
        std::string v;
        in >> h >> v;
        const double mine = num.cfl(P(h)), theirs = hexd(v);
        ++cflChecks;
        if (!(mine == theirs)) {
          if (cflMismatch++ < 5)
            std::fprintf(stderr, "cfl of patch %d: %a, expected %a\n", h, mine, theirs);
        }
      } else if (op == "K") {
        int h;
        std::string t, dt;
        in >> h >> t >> dt;
        auto& ti = P(h).getTimeIntervals();
        ++clockChecks;
        if (!(ti.getCurrentTime() == hexd(t) && ti.getTimestepSize() == hexd(dt))) {
          if (clockMismatch++ < 5)
            std::fprintf(stderr, "clock of patch %d: (%a, %a), expected (%s, %s)\n", h,
                         ti.getCurrentTime(), ti.getTimestepSize(), t.c_str(), dt.c_str());
        }
      } else if (op == "X") {
        int h;
        std::string m;
        in >> h >> m;
        const std::array<bool, 4> flags{false, false, false, false};
        num.solveTimestep(P(h), hexd(m), false, flags);
      } else if (op == "D") {
        int h;
        std::string thr, minw, maxw, w0, w1;
        in >> h >> thr >> minw >> maxw >> w0 >> w1;
        num.setRefinementCriterion(hexd(thr), hexd(minw), hexd(maxw));
        const std::pair<double, double> w = num.getDemandedMeshWidth(P(h), false);
        P(h).setDemandedMeshWidth(w.first, w.second);
        ++widthChecks;
        if (!(w.first == hexd(w0) && w.second == hexd(w1))) {
          if (widthMismatch++ < 5)
            std::fprintf(stderr, "demanded width of patch %d: (%a, %a), expected (%s, %s)\n", h,
                         w.first, w.second, w0.c_str(), w1.c_str());
        }
      } else if (op == "U" || op == "B") {
        int h;
        size_t off;
        in >> h >> off;
        HeapPatch& p = P(h);
        const int n = p.n();
        const size_t a = (size_t)n * n, g = (size_t)(n + 2) * (n + 2);
        std::vector<double>& u = p.heap();
        num.downloadToHeap(p);
        if (op == "U") {
          std::string t, dt;
          in >> t >> dt;
          if (off + 3 * g + 3 * a > blob.size())
            throw peanoclaw_amd::Error(PC_E_INVALID, "blob too short");
          std::copy(blob.begin() + off, blob.begin() + off + 3 * g, u.begin() + 3 * a);
          std::copy(blob.begin() + off + 3 * g, blob.begin() + off + 3 * g + 3 * a, u.begin());
          p.setVirtual(false);
          p.getTimeIntervals().setCurrentTime(hexd(t));
          p.getTimeIntervals().setTimestepSize(hexd(dt));
        } else {
          if (off + g > blob.size()) throw peanoclaw_amd::Error(PC_E_INVALID, "blob too short");
          std::copy(blob.begin() + off, blob.begin() + off + g, u.begin() + 3 * a + 3 * g);
        }
        num.uploadFromHeap(p);
      } else if (op == "Q") {
        int h;
        in >> h;
        num.destroyPatch(P(h));
        patches[h].reset();
      } else {
        throw peanoclaw_amd::Error(PC_E_INVALID, "unknown trace operation " + op);
      }
      if (!in) throw peanoclaw_amd::Error(PC_E_INVALID, "malformed trace line: " + line);
    }
    FILE* out = std::fopen(argv[3], "wb");
    if (!out) throw peanoclaw_amd::Error(PC_E_INVALID, std::string("cannot write ") + argv[3]);
    bool ok = true;
    for (auto& pp : patches) {
      if (!pp) {  // released: no record (the reader skips it)
        continue;
      }
      HeapPatch& p = *pp;
      num.addPatchToSolution(p);
      auto& ti = p.getTimeIntervals();
      const double clk[3] = {ti.getCurrentTime(), ti.getTimestepSize(),
                             ti.getEstimatedNextTimestepSize()};
      const size_t a = 3 * (size_t)p.n() * p.n();
      ok = ok && std::fwrite(clk, sizeof(double), 3, out) == 3;
      ok = ok && std::fwrite(p.heap().data(), sizeof(double), a, out) == a;
    }
    ok = (std::fclose(out) == 0) && ok;
    if (!ok) throw peanoclaw_amd::Error(PC_E_INVALID, "write failed");
  } catch (const peanoclaw_amd::Error& e) {
    std::fprintf(stderr, "error (status %d) after %ld calls: %s\n", e.code(), calls, e.what());
    return 2;
  }
  std::printf("replayed %ld calls on %zu patches; cfl %ld/%ld equal, clocks %ld/%ld equal, "
              "demanded widths %ld/%ld equal\n",
              calls, patches.size(), cflChecks - cflMismatch, cflChecks,
              clockChecks - clockMismatch, clockChecks, widthChecks - widthMismatch, widthChecks);
  return (cflMismatch || clockMismatch || widthMismatch) ? 1 : 0;
}
