// bm_tune.cpp -- autotuner of the kernel tunables, cached per GPU
// architecture and shape class in a JSON file.
//
// The tunables (Tuning) are the EDT tile heights (rows of the Y-pass tile and
// lines of the Z-pass tile, with their halos) and the fill's candidate
// capacity per LDS round (the ridge-bin size of a round).  They depend on the
// extents along y and z, so the shape class is the pair (H, D), each rounded
// up to a multiple of 64.  The tuner times every candidate on a real volume of
// the class with HIP events and keeps the fastest of each pass.  The results
// never depend on the tunables: every choice is exact, only the speed changes.
//
// Cache file:
//   {"gfx950": {"h384_d320": {"y_rows": 512, "y_halo": 64, "z_rows": 256,
//                             "z_halo": 32, "fill_cap": 1024, ...}, ...}}
// found at BM_TUNE_CACHE (environment) or given to bm_tuning_load().
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cctype>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <map>
#include <mutex>
#include <sstream>
#include <string>
#include <vector>

#include "bm_analyze.h"
#include "bm_engine.h"

namespace bm {

std::string shape_class(const Dims& dims) {
  auto r64 = [](uint32_t v) { return (v + 63) / 64 * 64; };
  char buf[64];
  snprintf(buf, sizeof buf, "h%u_d%u", r64(dims.h), r64(dims.d));
  return buf;
}

// Architecture of the current device ("gfx950"), looked up once per device.
static std::string device_arch() {
  static std::mutex mu;
  static std::map<int, std::string> cache;
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) return "unknown";
  std::lock_guard<std::mutex> g(mu);
  auto it = cache.find(dev);
  if (it != cache.end()) return it->second;
  hipDeviceProp_t prop;
  if (hipGetDeviceProperties(&prop, dev) != hipSuccess) return "unknown";
  std::string a = prop.gcnArchName;
  const size_t c = a.find(':');
  a = c == std::string::npos ? a : a.substr(0, c);
  cache[dev] = a;
  return a;
}

// ---------------------------------------------------------------------------
// A small reader for the cache's JSON subset: objects, strings, numbers.
namespace {
struct JsonReader {
  const std::string& s;
  size_t i = 0;
  explicit JsonReader(const std::string& text) : s(text) {}
  void ws() {
    while (i < s.size() && std::isspace((unsigned char)s[i])) ++i;
  }
  bool eat(char c) {
    ws();
    if (i < s.size() && s[i] == c) {
      ++i;
      return true;
    }
    return false;
  }
  std::string str() {
    ws();
    if (i >= s.size() || s[i] != '"') throw Error("tuning cache: string expected");
    ++i;
    std::string out;
    while (i < s.size() && s[i] != '"') {
      if (s[i] == '\\' && i + 1 < s.size()) ++i;
      out += s[i++];
    }
    ++i;
    return out;
  }
  double num() {
    ws();
    const char* p = s.c_str() + i;
    char* end = nullptr;
    const double v = std::strtod(p, &end);
    if (end == p) throw Error("tuning cache: number expected");
    i += (size_t)(end - p);
    return v;
  }
  // object of numbers
  std::map<std::string, double> flat() {
    std::map<std::string, double> m;
    if (!eat('{')) throw Error("tuning cache: object expected");
    if (eat('}')) return m;
    do {
      const std::string k = str();
      if (!eat(':')) throw Error("tuning cache: ':' expected");
      m[k] = num();
    } while (eat(','));
    if (!eat('}')) throw Error("tuning cache: '}' expected");
    return m;
  }
};

struct TuneEntry {
  Tuning t;
  std::map<std::string, double> extra;  // measured times, kept in the file
};

std::mutex g_tune_mu;
bool g_tune_loaded = false;
std::string g_tune_path;
// arch -> class -> entry
std::map<std::string, std::map<std::string, TuneEntry>> g_tune;

void parse_cache(const std::string& text) {
  JsonReader r(text);
  g_tune.clear();
  if (!r.eat('{')) throw Error("tuning cache: object expected");
  if (r.eat('}')) return;
  do {
    const std::string arch = r.str();
    if (!r.eat(':') || !r.eat('{')) throw Error("tuning cache: object expected");
    auto& classes = g_tune[arch];
    if (!r.eat('}')) {
      do {
        const std::string cls = r.str();
        if (!r.eat(':')) throw Error("tuning cache: ':' expected");
        const auto m = r.flat();
        TuneEntry e;
        auto get = [&](const char* k, uint32_t& dst) {
          auto it = m.find(k);
          if (it != m.end()) dst = (uint32_t)it->second;
        };
        get("y_rows", e.t.y_rows);
        get("y_halo", e.t.y_halo);
        get("z_rows", e.t.z_rows);
        get("z_halo", e.t.z_halo);
        get("fill_cap", e.t.fill_cap);
        get("tile_order", e.t.tile_order);
        get("fill_tiles", e.t.fill_tiles);
        get("cleanup_blocks", e.t.cleanup_blocks);
        for (const auto& kv : m)
          if (kv.first.find("_ms") != std::string::npos) e.extra[kv.first] = kv.second;
        classes[cls] = e;
      } while (r.eat(','));
      if (!r.eat('}')) throw Error("tuning cache: '}' expected");
    }
  } while (r.eat(','));
}

std::string dump_cache() {
  std::ostringstream o;
  o << "{\n";
  bool first_arch = true;
  for (const auto& a : g_tune) {
    o << (first_arch ? "" : ",\n") << "  \"" << a.first << "\": {\n";
    first_arch = false;
    bool first = true;
    for (const auto& c : a.second) {
      const Tuning& t = c.second.t;
      o << (first ? "" : ",\n") << "    \"" << c.first << "\": {\"y_rows\": " << t.y_rows
        << ", \"y_halo\": " << t.y_halo << ", \"z_rows\": " << t.z_rows
        << ", \"z_halo\": " << t.z_halo << ", \"fill_cap\": " << t.fill_cap
        << ", \"tile_order\": " << t.tile_order << ", \"fill_tiles\": " << t.fill_tiles;
      if (t.cleanup_blocks) o << ", \"cleanup_blocks\": " << t.cleanup_blocks;
      for (const auto& kv : c.second.extra) {
        char buf[64];
        snprintf(buf, sizeof buf, "%.4f", kv.second);
        o << ", \"" << kv.first << "\": " << buf;
      }
      o << "}";
      first = false;
    }
    o << "\n  }";
  }
  o << "\n}\n";
  return o.str();
}

void load_locked(const std::string& path) {
  std::ifstream f(path);
  if (!f) {  // no cache yet: defaults
    g_tune.clear();
    g_tune_path = path;
    g_tune_loaded = true;
    return;
  }
  std::stringstream ss;
  ss << f.rdbuf();
  parse_cache(ss.str());
  g_tune_path = path;
  g_tune_loaded = true;
}

void ensure_loaded() {
  if (g_tune_loaded) return;
  const char* env = std::getenv("BM_TUNE_CACHE");
  if (env && *env) {
    try {
      load_locked(env);
    } catch (const std::exception&) {
      g_tune.clear();  // an unreadable cache means defaults
    }
  }
  g_tune_loaded = true;
}
}  // namespace

Tuning tuned_for(const Dims& dims) {
  std::lock_guard<std::mutex> g(g_tune_mu);
  ensure_loaded();
  if (g_tune.empty()) return Tuning();
  auto a = g_tune.find(device_arch());
  if (a == g_tune.end()) return Tuning();
  auto c = a->second.find(shape_class(dims));
  return c == a->second.end() ? Tuning() : c->second.t;
}

void tuning_load(const std::string& path) {
  std::lock_guard<std::mutex> g(g_tune_mu);
  load_locked(path);
}

// ---------------------------------------------------------------------------
// The tuner.  For each sample (the largest volume of a class, device-resident
// bone and ROI) it times, on the marrow phase (the slower one), every
// candidate of each pass with the others at their defaults.
static double time_kernel(Engine& e, const ItemSpec& sp, const char* kernel, int reps) {
  Plan& pl = e.plain_plan();
  hipStream_t s = e.stream();
  pl.build({sp}, s);
  pl.launch(s);  // warm-up
  pl.fetch(s);
  KernelTime kt[64];
  profile_read(kt, 64);
  double best = 1e30;
  for (int r = 0; r < reps; ++r) {
    profile_enable(true);
    pl.launch(s);
    profile_enable(false);
    const int n = profile_read(kt, 64);
    for (int i = 0; i < n; ++i)
      if (std::strcmp(kt[i].name, kernel) == 0) best = std::min(best, kt[i].total_ms);
  }
  pl.fetch(s);
  return best;
}

void autotune(const std::vector<TuneSample>& samples, const std::string& path, bool verbose) {
  Engine& e = engine_for_device(current_device());
  std::lock_guard<std::mutex> g(e.mutex());
  const std::string arch = device_arch();
  // (rows, halo) of the EDT tiles; a split line needs a chunk of at least kMinZChunk
  const uint32_t tile_c[][2] = {{64, 8},   {128, 16}, {128, 24}, {128, 32}, {192, 16}, {192, 24},
                                {256, 32}, {256, 48}, {384, 16}, {384, 32}, {512, 64}};
  const uint32_t cap_c[] = {512, 1024};
  std::map<std::string, TuneEntry> found;
  for (const auto& smp : samples) {
    ItemSpec base;
    base.src.data = smp.bone;
    base.src.roi = smp.roi;
    base.src.want_nonzero = 0;  // marrow
    base.dims = smp.dims;
    base.thickness = false;
    base.keep_d2 = false;
    TuneEntry best;
    double by = 1e30, bz = 1e30, bf = 1e30;
    for (const auto& tc : tile_c) {
      const uint32_t rows = tc[0], halo = tc[1];
      // a split column needs a chunk; the whole column is one candidate
      if (rows < smp.dims.h || rows == 512) {
        ItemSpec sp = base;
        sp.tune.y_rows = rows;
        sp.tune.y_halo = halo;
        const double ms = time_kernel(e, sp, "edt_y", 3);
        if (verbose) fprintf(stderr, "[tune] %s y_rows %u halo %u: %.3f ms\n", shape_class(smp.dims).c_str(), rows, halo, ms);
        if (ms < by) {
          by = ms;
          best.t.y_rows = rows;
          best.t.y_halo = halo;
        }
      }
      if ((rows < smp.dims.d && rows - 2 * halo >= kMinZChunk) || rows == 512) {
        ItemSpec sp = base;
        sp.tune.z_rows = rows;
        sp.tune.z_halo = halo;
        const double ms = time_kernel(e, sp, "edt_z", 3);
        if (verbose) fprintf(stderr, "[tune] %s z_rows %u halo %u: %.3f ms\n", shape_class(smp.dims).c_str(), rows, halo, ms);
        if (ms < bz) {
          bz = ms;
          best.t.z_rows = rows;
          best.t.z_halo = halo;
        }
      }
    }
    for (uint32_t cap : cap_c) {
      ItemSpec sp = base;
      sp.thickness = true;
      sp.tune = best.t;
      sp.tune.fill_cap = cap;
      const double ms = time_kernel(e, sp, "fill_gather", 2);
      if (verbose) fprintf(stderr, "[tune] %s fill_cap %u: %.3f ms\n", shape_class(smp.dims).c_str(), cap, ms);
      if (ms < bf) {
        bf = ms;
        best.t.fill_cap = cap;
      }
    }
    for (uint32_t tiles : {2u}) {  // fill blocks of two tiles against the one-tile default
      ItemSpec sp = base;
      sp.thickness = true;
      sp.tune = best.t;
      sp.tune.fill_tiles = tiles;
      const double ms = time_kernel(e, sp, "fill_gather", 2);
      if (verbose) fprintf(stderr, "[tune] %s fill_tiles %u: %.3f ms\n", shape_class(smp.dims).c_str(), tiles, ms);
      if (ms < bf) {
        bf = ms;
        best.t.fill_tiles = tiles;
      }
    }
    best.extra["edt_y_ms"] = by;
    best.extra["edt_z_ms"] = bz;
    best.extra["fill_gather_ms"] = bf;
    found[shape_class(smp.dims)] = best;
  }
  std::lock_guard<std::mutex> g2(g_tune_mu);
  try {
    load_locked(path);  // merge into what the file already holds
  } catch (const std::exception&) {
    g_tune.clear();
  }
  // (cleanup_blocks is not swept: a class keeps the value its entry had)
  for (auto& kv : found) {
    TuneEntry& slot = g_tune[arch][kv.first];
    const uint32_t cleanup_blocks = slot.t.cleanup_blocks;
    slot = kv.second;
    slot.t.cleanup_blocks = cleanup_blocks;
  }
  std::ofstream f(path);
  if (!f) throw Error("cannot write the tuning cache " + path);
  f << dump_cache();
  g_tune_path = path;
  g_tune_loaded = true;
}

}  // namespace bm
