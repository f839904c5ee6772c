// Warm restart: snapshot keys and flag hashes, the device-resident exact
// snapshot, the compact macroscopic snapshot with its FCS1 file format, and
// the LRU disk cache of compact snapshot files (reference src/sim/snapshot.h,
// src/sim/snapshot.cpp).
//
// Exact snapshots are independent objects, each holding its own device copy:
// with 288 GB of HBM per GPU a caller keeps as many slots as it needs (one
// per polar point, a clean-foil slot plus VG variants) instead of the single
// slot a consumer card allows.
#include <hip/hip_runtime_api.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <deque>
#include <filesystem>
#include <fstream>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "../../include/foilcfd.h"

namespace fc {
// Bridge into the solver object (defined with it).
bool solverReady(fc_solver* s);
void solverStateRegions(fc_solver* s, std::vector<std::pair<void*, std::size_t>>& out, int layouts[3],
                        unsigned long long* roundSeq);
int solverStateRestored(fc_solver* s, const int layouts[3], unsigned long long roundSeq);
void* solverStream(fc_solver* s);
}  // namespace fc

namespace {

namespace fs = std::filesystem;

// FNV-1a 64 (snapshot.cpp:33-45): offset basis and prime of the published
// Fowler-Noll-Vo definition.
constexpr std::uint64_t kFnvOffset = 0xCBF29CE484222325ull;
constexpr std::uint64_t kFnvPrime = 0x100000001B3ull;

std::uint64_t fnv1a(const void* data, std::size_t n, std::uint64_t h = kFnvOffset) {
    const auto* p = static_cast<const std::uint8_t*>(data);
    for (std::size_t i = 0; i < n; ++i) {
        h ^= p[i];
        h *= kFnvPrime;
    }
    return h;
}

// FCS1 container (snapshot.cpp:47-60): magic "FCS1", container revision 1,
// payload compression 0 = raw FP32 arrays. All fields little-endian, written
// one by one (no struct images).
constexpr std::uint32_t kFileMagic = 0x31534346u;
constexpr std::uint32_t kFileFormat = 1;
constexpr std::uint32_t kCompressionRaw = 0;
constexpr std::uint32_t kMaxIdLen = 4096;

struct Key {
    std::string id;
    int aoaBucket = 0;
    float uLat = 0.0f;
    int nx = 0, ny = 0, nz = 0;
    int version = FC_SOLVER_VERSION;

    static Key from(const fc_snapshot_key& k) {
        Key o;
        o.id = k.geometry_id ? k.geometry_id : "";
        o.aoaBucket = k.aoa_bucket;
        o.uLat = k.u_lat;
        o.nx = k.dims.nx;
        o.ny = k.dims.ny;
        o.nz = k.dims.nz;
        o.version = k.solver_version;
        return o;
    }
    long long cells() const { return static_cast<long long>(nx) * ny * nz; }
    bool sameGrid(const fc_dims& d) const { return d.nx == nx && d.ny == ny && d.nz == nz; }

    // SnapshotKey::hash (snapshot.cpp:89-99): the id bytes, then each field's
    // native bytes, chained.
    std::uint64_t hash() const {
        std::uint64_t h = fnv1a(id.data(), id.size());
        h = fnv1a(&aoaBucket, sizeof aoaBucket, h);
        h = fnv1a(&uLat, sizeof uLat, h);
        h = fnv1a(&nx, sizeof nx, h);
        h = fnv1a(&ny, sizeof ny, h);
        h = fnv1a(&nz, sizeof nz, h);
        h = fnv1a(&version, sizeof version, h);
        return h;
    }

    // SnapshotKey::toString (snapshot.cpp:101-112).
    std::string str() const {
        char buf[256];
        std::snprintf(buf, sizeof buf, "%s_a%.1f_u%.4f_%dx%dx%d_v%d", id.c_str(),
                      static_cast<double>(static_cast<float>(aoaBucket) * 0.5f), static_cast<double>(uLat), nx,
                      ny, nz, version);
        std::string s(buf);
        for (char& c : s)
            if (c == ':' || c == '\\' || c == '/') c = '-';
        return s;
    }

    std::string fileName() const {
        char buf[32];
        std::snprintf(buf, sizeof buf, "%016llx.fcs", static_cast<unsigned long long>(hash()));
        return buf;
    }
};

int exportKey(const Key& k, fc_snapshot_key* out, char* idBuf, int idCap) {
    if (!out) return FC_ERR_INVALID;
    if (idBuf) {
        if (idCap <= static_cast<int>(k.id.size())) return FC_ERR_INVALID;
        std::memcpy(idBuf, k.id.data(), k.id.size());
        idBuf[k.id.size()] = '\0';
    }
    out->geometry_id = idBuf;
    out->aoa_bucket = k.aoaBucket;
    out->u_lat = k.uLat;
    out->dims = fc_dims{k.nx, k.ny, k.nz};
    out->solver_version = k.version;
    return FC_OK;
}

void setErr(const std::string& msg, char* err, int errlen) {
    if (!err || errlen <= 0) return;
    const std::size_t n = std::min(msg.size(), static_cast<std::size_t>(errlen - 1));
    std::memcpy(err, msg.data(), n);
    err[n] = '\0';
}

std::uint64_t hashSolverFlags(fc_solver* s, const fc_dims& d, int& rc) {
    std::vector<std::uint8_t> flags(static_cast<std::size_t>(d.nx) * d.ny * d.nz);
    rc = fc_read_flags(s, flags.data());
    return fnv1a(flags.data(), flags.size());
}

// ---- compact snapshot data (immutable once built) ---------------------------
struct Compact {
    Key key;
    std::uint64_t flagsHash = 0;
    fc_scaling scaling{};
    std::vector<float> rho, u, v, w;
};

template <class T>
void put(std::ofstream& os, const T& v) {
    os.write(reinterpret_cast<const char*>(&v), sizeof(T));
}
template <class T>
bool get(std::ifstream& is, T& v) {
    is.read(reinterpret_cast<char*>(&v), sizeof(T));
    return is.good();
}

// CompactSnapshot::saveToFile (snapshot.cpp:287-347).
bool saveCompact(const Compact& c, const fs::path& path, std::string& why) {
    std::ofstream os(path, std::ios::binary | std::ios::trunc);
    if (!os) {
        why = "cannot open " + path.string() + " for writing";
        return false;
    }
    put(os, kFileMagic);
    put(os, kFileFormat);
    const auto idLen = static_cast<std::uint32_t>(c.key.id.size());
    put(os, idLen);
    os.write(c.key.id.data(), idLen);
    put(os, c.key.aoaBucket);
    put(os, c.key.uLat);
    put(os, c.key.nx);
    put(os, c.key.ny);
    put(os, c.key.nz);
    put(os, c.key.version);
    put(os, c.flagsHash);
    const fc_scaling& s = c.scaling;
    put(os, s.chord_m);
    put(os, s.airspeed_ms);
    put(os, s.nu_m2s);
    put(os, s.chord_cells);
    put(os, s.u_lat);
    put(os, s.dx);
    put(os, s.dt);
    put(os, s.nu_lat_target);
    put(os, s.nu_lat);
    put(os, s.tau);
    put(os, static_cast<std::uint8_t>(s.tau_clamped ? 1 : 0));
    put(os, s.re_target);
    put(os, s.re_effective);
    put(os, kCompressionRaw);
    const std::uint64_t n = c.rho.size();
    put(os, n);
    for (const std::vector<float>* a : {&c.rho, &c.u, &c.v, &c.w})
        os.write(reinterpret_cast<const char*>(a->data()), static_cast<std::streamsize>(a->size() * sizeof(float)));
    os.flush();
    if (!os.good()) {
        why = "write failed (disk full?) at " + path.string();
        return false;
    }
    return true;
}

// CompactSnapshot::loadFromFile (snapshot.cpp:349-424): the same rejections.
bool loadCompact(Compact& c, const fs::path& path, std::string& why) {
    std::ifstream is(path, std::ios::binary);
    if (!is) {
        why = "cannot open " + path.string();
        return false;
    }
    auto reject = [&](const char* what) {
        why = std::string(what) + ": " + path.string();
        return false;
    };
    std::uint32_t magic = 0, fmt = 0;
    if (!get(is, magic) || magic != kFileMagic) return reject("bad magic");
    if (!get(is, fmt) || fmt != kFileFormat) return reject("unsupported format version");
    std::uint32_t idLen = 0;
    if (!get(is, idLen) || idLen > kMaxIdLen) return reject("corrupt geometry id");
    c.key.id.resize(idLen);
    is.read(c.key.id.data(), idLen);
    if (!get(is, c.key.aoaBucket) || !get(is, c.key.uLat) || !get(is, c.key.nx) || !get(is, c.key.ny) ||
        !get(is, c.key.nz) || !get(is, c.key.version))
        return reject("truncated header");
    if (c.key.version != FC_SOLVER_VERSION) return reject("solver version mismatch");
    if (!get(is, c.flagsHash)) return reject("truncated header");
    fc_scaling& s = c.scaling;
    std::uint8_t clamped = 0;
    const bool ok = get(is, s.chord_m) && get(is, s.airspeed_ms) && get(is, s.nu_m2s) && get(is, s.chord_cells) &&
                    get(is, s.u_lat) && get(is, s.dx) && get(is, s.dt) && get(is, s.nu_lat_target) &&
                    get(is, s.nu_lat) && get(is, s.tau) && get(is, clamped) && get(is, s.re_target) &&
                    get(is, s.re_effective);
    if (!ok) return reject("truncated scaling block");
    s.tau_clamped = clamped != 0 ? 1 : 0;
    std::uint32_t compression = 0;
    std::uint64_t n = 0;
    if (!get(is, compression) || compression != kCompressionRaw) return reject("unsupported payload compression");
    if (!get(is, n)) return reject("truncated payload header");
    const long long expect = c.key.cells();
    if (n == 0 || expect <= 0 || n != static_cast<std::uint64_t>(expect)) return reject("payload size mismatch");
    for (std::vector<float>* a : {&c.rho, &c.u, &c.v, &c.w}) {
        a->resize(static_cast<std::size_t>(n));
        is.read(reinterpret_cast<char*>(a->data()), static_cast<std::streamsize>(n * sizeof(float)));
        if (!is.good()) return reject("truncated field payload");
    }
    return true;
}

}  // namespace

// ---- exact (device-resident) snapshot ----------------------------------------
struct fc_vram_snapshot {
    Key key;
    std::uint64_t flagsHash = 0;
    bool hasData = false;
    void* d = nullptr;
    std::size_t bytes = 0;
    std::vector<std::size_t> sizes;  // region sizes at capture (the grid's byte shape)
    int layouts[3] = {0, -1, -1};
    unsigned long long roundSeq = 0;  // the solver's rounding sequence position
    int storage = -1;
    fc_dims dims{0, 0, 0};

    void release() {
        if (d) (void)hipFree(d);
        d = nullptr;
        bytes = 0;
        sizes.clear();
        hasData = false;
    }
};

// ---- compact snapshot handle -------------------------------------------------
struct fc_compact_snapshot {
    std::shared_ptr<const Compact> data;
};

// ---- LRU disk cache (snapshot.cpp:431-611) -----------------------------------
struct fc_snapshot_cache {
    fs::path dir;
    std::uint64_t budget = 0;
    mutable std::mutex m;
    std::uint64_t used = 0;

    std::thread worker;
    std::mutex qm;
    std::condition_variable qcv, idle;
    std::deque<std::shared_ptr<const Compact>> queue;
    bool busy = false;
    bool stop = false;
    long long asyncFailures = 0;

    void scanLocked() {
        used = 0;
        std::error_code ec;
        for (const auto& e : fs::directory_iterator(dir, ec)) {
            if (!e.is_regular_file(ec) || e.path().extension() != ".fcs") continue;
            used += static_cast<std::uint64_t>(e.file_size(ec));
        }
    }

    // Oldest modification time first until the total fits; `keep` (the entry
    // just stored) is never evicted.
    void evictLocked(const fs::path& keep) {
        while (used > budget) {
            std::error_code ec;
            fs::path oldest;
            auto oldestTime = fs::file_time_type::max();
            std::uint64_t oldestSize = 0;
            for (const auto& e : fs::directory_iterator(dir, ec)) {
                if (!e.is_regular_file(ec) || e.path().extension() != ".fcs" || e.path() == keep) continue;
                const auto t = e.last_write_time(ec);
                if (t < oldestTime) {
                    oldestTime = t;
                    oldest = e.path();
                    oldestSize = static_cast<std::uint64_t>(e.file_size(ec));
                }
            }
            if (oldest.empty()) break;
            if (!fs::remove(oldest, ec)) {
                std::fprintf(stderr, "[cache] eviction of %s failed\n", oldest.string().c_str());
                break;
            }
            used = used > oldestSize ? used - oldestSize : 0;
        }
    }

    bool store(const Compact& c, std::string& why) {
        std::lock_guard<std::mutex> lk(m);
        const std::string name = c.key.fileName();
        const fs::path finalPath = dir / name;
        const fs::path tmpPath = dir / (name + ".tmp");
        std::error_code ec;
        // Temp file then rename: a crash mid-write never leaves a truncated .fcs.
        if (!saveCompact(c, tmpPath, why)) {
            fs::remove(tmpPath, ec);
            return false;
        }
        std::uint64_t replaced = 0;
        if (fs::exists(finalPath, ec)) replaced = static_cast<std::uint64_t>(fs::file_size(finalPath, ec));
        fs::rename(tmpPath, finalPath, ec);
        if (ec) {
            why = "rename to " + finalPath.string() + " failed: " + ec.message();
            fs::remove(tmpPath, ec);
            return false;
        }
        const auto added = static_cast<std::uint64_t>(fs::file_size(finalPath, ec));
        used = used - std::min(used, replaced) + added;
        evictLocked(finalPath);
        return true;
    }

    void workerLoop() {
        for (;;) {
            std::shared_ptr<const Compact> snap;
            {
                std::unique_lock<std::mutex> lk(qm);
                qcv.wait(lk, [this] { return stop || !queue.empty(); });
                if (queue.empty()) return;
                snap = std::move(queue.front());
                queue.pop_front();
                busy = true;
            }
            std::string why;
            const bool ok = store(*snap, why);
            if (!ok) std::fprintf(stderr, "[cache] background store failed: %s\n", why.c_str());
            {
                std::lock_guard<std::mutex> lk(qm);
                busy = false;
                if (!ok) ++asyncFailures;
            }
            idle.notify_all();
        }
    }
};

extern "C" {

// ---- keys --------------------------------------------------------------------
int fc_snapshot_aoa_bucket(float aoa_deg) { return static_cast<int>(std::lround(aoa_deg * 2.0f)); }

uint64_t fc_snapshot_key_hash(const fc_snapshot_key* key) { return key ? Key::from(*key).hash() : 0; }

int fc_snapshot_key_string(const fc_snapshot_key* key, char* out, int cap) {
    if (!key || !out || cap <= 0) return FC_ERR_INVALID;
    const std::string s = Key::from(*key).str();
    if (cap <= static_cast<int>(s.size())) return FC_ERR_INVALID;
    std::memcpy(out, s.c_str(), s.size() + 1);
    return FC_OK;
}

int fc_snapshot_key_file_name(const fc_snapshot_key* key, char* out, int cap) {
    if (!key || !out || cap <= 0) return FC_ERR_INVALID;
    const std::string s = Key::from(*key).fileName();
    if (cap <= static_cast<int>(s.size())) return FC_ERR_INVALID;
    std::memcpy(out, s.c_str(), s.size() + 1);
    return FC_OK;
}

int fc_snapshot_key_equal(const fc_snapshot_key* a, const fc_snapshot_key* b) {
    if (!a || !b) return 0;
    const Key x = Key::from(*a), y = Key::from(*b);
    return (x.id == y.id && x.aoaBucket == y.aoaBucket && x.uLat == y.uLat && x.nx == y.nx && x.ny == y.ny &&
            x.nz == y.nz && x.version == y.version)
               ? 1
               : 0;
}

uint64_t fc_hash_flags(const uint8_t* flags, uint64_t count) {
    return flags || count == 0 ? fnv1a(flags, static_cast<std::size_t>(count)) : 0;
}

// ---- exact snapshot ----------------------------------------------------------
fc_vram_snapshot* fc_vram_snapshot_create(void) { return new (std::nothrow) fc_vram_snapshot(); }

void fc_vram_snapshot_destroy(fc_vram_snapshot* s) {
    if (!s) return;
    s->release();
    delete s;
}

int fc_vram_snapshot_capture(fc_vram_snapshot* snap, fc_solver* s, const fc_snapshot_key* key, char* err,
                             int errlen) {
    if (!snap || !key) return FC_ERR_INVALID;
    if (!fc::solverReady(s)) {
        setErr("solver holds no field to capture", err, errlen);
        return FC_ERR_STATE;
    }
    std::vector<std::pair<void*, std::size_t>> regions;
    int layouts[3];
    unsigned long long roundSeq = 0;
    fc::solverStateRegions(s, regions, layouts, &roundSeq);
    std::vector<std::size_t> sizes;
    std::size_t need = 0;
    for (const auto& r : regions) {
        sizes.push_back(r.second);
        need += (r.second + 255) & ~static_cast<std::size_t>(255);
    }
    // Re-captures of the same byte shape reuse the buffer.
    if (!snap->d || snap->bytes != need) {
        snap->release();
        if (hipMalloc(&snap->d, need) != hipSuccess) {
            (void)hipGetLastError();
            snap->d = nullptr;
            setErr("device snapshot allocation of " + std::to_string(need) + " bytes failed", err, errlen);
            return FC_ERR_ALLOC;
        }
        snap->bytes = need;
    }
    snap->hasData = false;
    const auto st = static_cast<hipStream_t>(fc::solverStream(s));
    std::size_t off = 0;
    for (const auto& r : regions) {
        if (hipMemcpyAsync(static_cast<char*>(snap->d) + off, r.first, r.second, hipMemcpyDeviceToDevice, st) !=
            hipSuccess) {
            setErr("device snapshot copy failed", err, errlen);
            return FC_ERR_DEVICE;
        }
        off += (r.second + 255) & ~static_cast<std::size_t>(255);
    }
    fc_get_dims(s, &snap->dims);
    int rc = FC_OK;
    snap->flagsHash = hashSolverFlags(s, snap->dims, rc);
    if (rc != FC_OK) {
        setErr("flag readback failed", err, errlen);
        return rc;
    }
    snap->key = Key::from(*key);
    snap->sizes = sizes;
    std::copy(layouts, layouts + 3, snap->layouts);
    snap->roundSeq = roundSeq;
    snap->storage = fc_storage(s);
    snap->hasData = true;
    return FC_OK;
}

int fc_vram_snapshot_restore(fc_vram_snapshot* snap, fc_solver* s) {
    if (!snap || !snap->hasData) return FC_ERR_STATE;
    if (!fc::solverReady(s)) return FC_ERR_STATE;
    fc_dims d{};
    fc_get_dims(s, &d);
    std::vector<std::pair<void*, std::size_t>> regions;
    int layouts[3];
    unsigned long long roundSeq = 0;
    fc::solverStateRegions(s, regions, layouts, &roundSeq);
    // The grid must be byte-identical (same dims, storage and level shapes).
    if (!snap->key.sameGrid(d) || d.nx != snap->dims.nx || d.ny != snap->dims.ny || d.nz != snap->dims.nz ||
        fc_storage(s) != snap->storage || regions.size() != snap->sizes.size())
        return FC_ERR_INVALID;
    for (std::size_t i = 0; i < regions.size(); ++i)
        if (regions[i].second != snap->sizes[i]) return FC_ERR_INVALID;
    const auto st = static_cast<hipStream_t>(fc::solverStream(s));
    std::size_t off = 0;
    for (const auto& r : regions) {
        if (hipMemcpyAsync(r.first, static_cast<const char*>(snap->d) + off, r.second, hipMemcpyDeviceToDevice,
                           st) != hipSuccess)
            return FC_ERR_DEVICE;
        off += (r.second + 255) & ~static_cast<std::size_t>(255);
    }
    return fc::solverStateRestored(s, snap->layouts, snap->roundSeq);
}

int fc_vram_snapshot_has_data(const fc_vram_snapshot* s) { return s && s->hasData ? 1 : 0; }

int fc_vram_snapshot_get_key(const fc_vram_snapshot* s, fc_snapshot_key* out, char* id_buf, int id_cap) {
    if (!s || !s->hasData) return FC_ERR_STATE;
    return exportKey(s->key, out, id_buf, id_cap);
}

uint64_t fc_vram_snapshot_flags_hash(const fc_vram_snapshot* s) { return s ? s->flagsHash : 0; }

long long fc_vram_snapshot_bytes(const fc_vram_snapshot* s) { return s ? static_cast<long long>(s->bytes) : 0; }

void fc_vram_snapshot_release(fc_vram_snapshot* s) {
    if (s) s->release();
}

// ---- compact snapshot --------------------------------------------------------
fc_compact_snapshot* fc_compact_snapshot_create(void) { return new (std::nothrow) fc_compact_snapshot(); }

void fc_compact_snapshot_destroy(fc_compact_snapshot* s) { delete s; }

int fc_compact_snapshot_capture(fc_compact_snapshot* snap, fc_solver* s, const fc_snapshot_key* key, char* err,
                                int errlen) {
    if (!snap || !key) return FC_ERR_INVALID;
    if (!fc::solverReady(s)) {
        setErr("solver holds no field to capture", err, errlen);
        return FC_ERR_STATE;
    }
    auto c = std::make_shared<Compact>();
    fc_dims d{};
    fc_get_dims(s, &d);
    const auto n = static_cast<std::size_t>(d.nx) * d.ny * d.nz;
    c->rho.resize(n);
    c->u.resize(n);
    c->v.resize(n);
    c->w.resize(n);
    // The macro arrays of the last stepped batch: the state on screen.
    if (fc_read_macro(s, c->rho.data(), c->u.data(), c->v.data(), c->w.data()) != FC_OK) {
        setErr("macroscopic download failed", err, errlen);
        return FC_ERR_DEVICE;
    }
    int rc = FC_OK;
    c->flagsHash = hashSolverFlags(s, d, rc);
    if (rc != FC_OK) {
        setErr("flag readback failed", err, errlen);
        return rc;
    }
    fc_get_scaling(s, &c->scaling);
    c->key = Key::from(*key);
    snap->data = std::move(c);
    return FC_OK;
}

int fc_compact_snapshot_set(fc_compact_snapshot* snap, const fc_snapshot_key* key, uint64_t flags_hash,
                            const fc_scaling* scaling, const float* rho, const float* u, const float* v,
                            const float* w, long long ncells) {
    if (!snap || !key || !scaling || !rho || !u || !v || !w) return FC_ERR_INVALID;
    auto c = std::make_shared<Compact>();
    c->key = Key::from(*key);
    if (ncells <= 0 || ncells != c->key.cells()) return FC_ERR_INVALID;
    c->flagsHash = flags_hash;
    c->scaling = *scaling;
    const auto n = static_cast<std::size_t>(ncells);
    c->rho.assign(rho, rho + n);
    c->u.assign(u, u + n);
    c->v.assign(v, v + n);
    c->w.assign(w, w + n);
    snap->data = std::move(c);
    return FC_OK;
}

int fc_compact_snapshot_restore(fc_compact_snapshot* snap, fc_solver* s) {
    if (!snap || !snap->data) return FC_ERR_STATE;
    if (!fc::solverReady(s)) return FC_ERR_STATE;
    fc_dims d{};
    fc_get_dims(s, &d);
    const Compact& c = *snap->data;
    if (!c.key.sameGrid(d)) return FC_ERR_INVALID;
    return fc_restore_from_macroscopic(s, c.rho.data(), c.u.data(), c.v.data(), c.w.data());
}

int fc_compact_snapshot_save(const fc_compact_snapshot* snap, const char* path, char* err, int errlen) {
    if (!snap || !path) return FC_ERR_INVALID;
    if (!snap->data) {
        setErr("no snapshot data to save", err, errlen);
        return FC_ERR_STATE;
    }
    std::string why;
    if (!saveCompact(*snap->data, fs::path(path), why)) {
        setErr(why, err, errlen);
        return FC_ERR_IO;
    }
    return FC_OK;
}

int fc_compact_snapshot_load(fc_compact_snapshot* snap, const char* path, char* err, int errlen) {
    if (!snap || !path) return FC_ERR_INVALID;
    snap->data.reset();
    auto c = std::make_shared<Compact>();
    std::string why;
    if (!loadCompact(*c, fs::path(path), why)) {
        setErr(why, err, errlen);
        return FC_ERR_IO;
    }
    snap->data = std::move(c);
    return FC_OK;
}

int fc_compact_snapshot_has_data(const fc_compact_snapshot* s) { return s && s->data ? 1 : 0; }

int fc_compact_snapshot_get_key(const fc_compact_snapshot* s, fc_snapshot_key* out, char* id_buf, int id_cap) {
    if (!s || !s->data) return FC_ERR_STATE;
    return exportKey(s->data->key, out, id_buf, id_cap);
}

uint64_t fc_compact_snapshot_flags_hash(const fc_compact_snapshot* s) {
    return s && s->data ? s->data->flagsHash : 0;
}

int fc_compact_snapshot_scaling(const fc_compact_snapshot* s, fc_scaling* out) {
    if (!s || !s->data || !out) return FC_ERR_STATE;
    *out = s->data->scaling;
    return FC_OK;
}

int fc_compact_snapshot_fields(const fc_compact_snapshot* s, float* rho, float* u, float* v, float* w) {
    if (!s || !s->data) return FC_ERR_STATE;
    const Compact& c = *s->data;
    const std::pair<float*, const std::vector<float>*> outs[] = {{rho, &c.rho}, {u, &c.u}, {v, &c.v}, {w, &c.w}};
    for (const auto& o : outs)
        if (o.first) std::memcpy(o.first, o.second->data(), o.second->size() * sizeof(float));
    return FC_OK;
}

// ---- disk cache --------------------------------------------------------------
fc_snapshot_cache* fc_snapshot_cache_open(const char* dir, uint64_t budget_bytes) {
    if (!dir) return nullptr;
    auto* c = new (std::nothrow) fc_snapshot_cache();
    if (!c) return nullptr;
    c->dir = fs::path(dir);
    c->budget = budget_bytes;
    std::error_code ec;
    fs::create_directories(c->dir, ec);
    if (!fs::is_directory(c->dir, ec)) {
        delete c;
        return nullptr;
    }
    std::lock_guard<std::mutex> lk(c->m);
    c->scanLocked();
    return c;
}

void fc_snapshot_cache_close(fc_snapshot_cache* c) {
    if (!c) return;
    {
        std::lock_guard<std::mutex> lk(c->qm);
        c->stop = true;
    }
    c->qcv.notify_all();
    if (c->worker.joinable()) c->worker.join();  // queued writes finish first
    delete c;
}

int fc_snapshot_cache_find(fc_snapshot_cache* c, const fc_snapshot_key* key, char* path_out, int cap) {
    if (!c || !key) return -1;
    std::lock_guard<std::mutex> lk(c->m);
    const fs::path p = c->dir / Key::from(*key).fileName();
    std::error_code ec;
    if (!fs::exists(p, ec)) return 0;
    fs::last_write_time(p, fs::file_time_type::clock::now(), ec);  // a hit refreshes recency
    if (path_out) {
        const std::string s = p.string();
        if (cap <= static_cast<int>(s.size())) return -1;
        std::memcpy(path_out, s.c_str(), s.size() + 1);
    }
    return 1;
}

int fc_snapshot_cache_store(fc_snapshot_cache* c, const fc_compact_snapshot* snap, char* err, int errlen) {
    if (!c || !snap) return FC_ERR_INVALID;
    if (!snap->data) {
        setErr("snapshot holds no data", err, errlen);
        return FC_ERR_STATE;
    }
    std::string why;
    if (!c->store(*snap->data, why)) {
        setErr(why, err, errlen);
        return FC_ERR_IO;
    }
    return FC_OK;
}

int fc_snapshot_cache_store_async(fc_snapshot_cache* c, const fc_compact_snapshot* snap) {
    if (!c || !snap || !snap->data) return FC_ERR_INVALID;
    {
        std::lock_guard<std::mutex> lk(c->qm);
        if (c->stop) return FC_ERR_STATE;
        if (!c->worker.joinable()) c->worker = std::thread([c] { c->workerLoop(); });
        // Shares the immutable data: later captures into the handle do not
        // disturb the pending write.
        c->queue.push_back(snap->data);
    }
    c->qcv.notify_one();
    return FC_OK;
}

long long fc_snapshot_cache_flush(fc_snapshot_cache* c) {
    if (!c) return -1;
    std::unique_lock<std::mutex> lk(c->qm);
    c->idle.wait(lk, [c] { return c->queue.empty() && !c->busy; });
    return c->asyncFailures;
}

int fc_snapshot_cache_set_budget(fc_snapshot_cache* c, uint64_t budget_bytes) {
    if (!c) return FC_ERR_INVALID;
    std::lock_guard<std::mutex> lk(c->m);
    c->budget = budget_bytes;
    c->evictLocked({});
    return FC_OK;
}

uint64_t fc_snapshot_cache_used_bytes(const fc_snapshot_cache* c) {
    if (!c) return 0;
    std::lock_guard<std::mutex> lk(c->m);
    return c->used;
}

}  // extern "C"
