// HDF5 project files (C-ABI in include/dxmc_h5.h) on the HDF5 C library,
// opened with dlopen at first use so that nothing else in libdxmc_hip.so
// depends on it, plus a multi-threaded deflate for compressed datasets.
//
// Compressed datasets are stored as the reference stores them (one chunk of
// the whole extent, deflate filter): the chunk is deflated here by up to 16
// threads into a single zlib stream -- 1 MiB blocks, each primed with the 32
// KiB before it as dictionary and ended by a sync flush, so the concatenation
// is one valid deflate stream; Adler-32 of the whole comes from the blocks'
// checksums combined -- and written with H5Dwrite_chunk as already-filtered
// data.  HDF5's deflate filter inflates it like any other.
#include "dxmc_h5.h"

#include "physics.hpp"

#include <hdf5.h>
#include <zlib.h>

#include <dlfcn.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

namespace {

int fail(int code, const std::string& msg) { return dxg::setError(code, msg); }

// H5F_ACC_RDONLY / H5F_ACC_TRUNC without the library calls the header's
// macros wrap around them (nothing here links the library).
constexpr unsigned kAccRdonly = 0x0000u;
constexpr unsigned kAccTrunc = 0x0002u;

// ---------------------------------------------------------------- library

struct Api {
    void* handle = nullptr;
#define DXG_H5_FN(name) decltype(&::name) name = nullptr;
    DXG_H5_FN(H5open)
    DXG_H5_FN(H5Eset_auto2)
    DXG_H5_FN(H5Fcreate)
    DXG_H5_FN(H5Fopen)
    DXG_H5_FN(H5Fclose)
    DXG_H5_FN(H5Gcreate2)
    DXG_H5_FN(H5Gopen2)
    DXG_H5_FN(H5Gclose)
    DXG_H5_FN(H5Lexists)
    DXG_H5_FN(H5Oopen)
    DXG_H5_FN(H5Oclose)
    DXG_H5_FN(H5Iget_type)
    DXG_H5_FN(H5Screate)
    DXG_H5_FN(H5Screate_simple)
    DXG_H5_FN(H5Sclose)
    DXG_H5_FN(H5Sget_simple_extent_ndims)
    DXG_H5_FN(H5Sget_simple_extent_dims)
    DXG_H5_FN(H5Sget_simple_extent_npoints)
    DXG_H5_FN(H5Pcreate)
    DXG_H5_FN(H5Pset_chunk)
    DXG_H5_FN(H5Pset_deflate)
    DXG_H5_FN(H5Pclose)
    DXG_H5_FN(H5Dcreate2)
    DXG_H5_FN(H5Dopen2)
    DXG_H5_FN(H5Dclose)
    DXG_H5_FN(H5Dwrite)
    DXG_H5_FN(H5Dread)
    DXG_H5_FN(H5Dget_space)
    DXG_H5_FN(H5Dget_type)
    DXG_H5_FN(H5Dwrite_chunk)
    DXG_H5_FN(H5Dvlen_reclaim)
    DXG_H5_FN(H5Tcopy)
    DXG_H5_FN(H5Tset_size)
    DXG_H5_FN(H5Tget_size)
    DXG_H5_FN(H5Tget_class)
    DXG_H5_FN(H5Tis_variable_str)
    DXG_H5_FN(H5Tclose)
    DXG_H5_FN(H5Acreate2)
    DXG_H5_FN(H5Aopen)
    DXG_H5_FN(H5Aexists)
    DXG_H5_FN(H5Awrite)
    DXG_H5_FN(H5Aread)
    DXG_H5_FN(H5Aget_space)
    DXG_H5_FN(H5Aget_type)
    DXG_H5_FN(H5Aclose)
    DXG_H5_FN(H5Literate)
#undef DXG_H5_FN
    // values of the library's predefined-type and property-class globals
    hid_t f64 = -1, u64 = -1, u8 = -1, f32 = -1, i32 = -1, i64 = -1, cs1 = -1, dcreate = -1;
};

std::mutex g_mutex;   // every HDF5 call below runs under it

const Api* loadApi(std::string& err)
{
    static Api api;
    static std::string load_err;
    static std::once_flag once;
    std::call_once(once, [] {
        std::vector<std::string> names;
        if (const char* e = std::getenv("DXG_HDF5_LIBRARY"))
            names.emplace_back(e);
        names.emplace_back("/opt/conda/lib/libhdf5.so.103");
        names.emplace_back("libhdf5.so");
        for (const auto& n : names) {
            api.handle = dlopen(n.c_str(), RTLD_NOW | RTLD_LOCAL);
            if (api.handle)
                break;
        }
        if (!api.handle) {
            load_err = "HDF5 library not found (set DXG_HDF5_LIBRARY to libhdf5.so)";
            return;
        }
        bool ok = true;
        auto sym = [&](const char* name) {
            void* p = dlsym(api.handle, name);
            if (!p) {
                ok = false;
                load_err = std::string("HDF5 library lacks ") + name;
            }
            return p;
        };
#define DXG_H5_LOAD(name) api.name = reinterpret_cast<decltype(api.name)>(sym(#name));
        DXG_H5_LOAD(H5open)
        DXG_H5_LOAD(H5Eset_auto2)
        DXG_H5_LOAD(H5Fcreate)
        DXG_H5_LOAD(H5Fopen)
        DXG_H5_LOAD(H5Fclose)
        DXG_H5_LOAD(H5Gcreate2)
        DXG_H5_LOAD(H5Gopen2)
        DXG_H5_LOAD(H5Gclose)
        DXG_H5_LOAD(H5Lexists)
        DXG_H5_LOAD(H5Oopen)
        DXG_H5_LOAD(H5Oclose)
        DXG_H5_LOAD(H5Iget_type)
        DXG_H5_LOAD(H5Screate)
        DXG_H5_LOAD(H5Screate_simple)
        DXG_H5_LOAD(H5Sclose)
        DXG_H5_LOAD(H5Sget_simple_extent_ndims)
        DXG_H5_LOAD(H5Sget_simple_extent_dims)
        DXG_H5_LOAD(H5Sget_simple_extent_npoints)
        DXG_H5_LOAD(H5Pcreate)
        DXG_H5_LOAD(H5Pset_chunk)
        DXG_H5_LOAD(H5Pset_deflate)
        DXG_H5_LOAD(H5Pclose)
        DXG_H5_LOAD(H5Dcreate2)
        DXG_H5_LOAD(H5Dopen2)
        DXG_H5_LOAD(H5Dclose)
        DXG_H5_LOAD(H5Dwrite)
        DXG_H5_LOAD(H5Dread)
        DXG_H5_LOAD(H5Dget_space)
        DXG_H5_LOAD(H5Dget_type)
        DXG_H5_LOAD(H5Tcopy)
        DXG_H5_LOAD(H5Tset_size)
        DXG_H5_LOAD(H5Tget_size)
        DXG_H5_LOAD(H5Tget_class)
        DXG_H5_LOAD(H5Tis_variable_str)
        DXG_H5_LOAD(H5Tclose)
        DXG_H5_LOAD(H5Acreate2)
        DXG_H5_LOAD(H5Aopen)
        DXG_H5_LOAD(H5Aexists)
        DXG_H5_LOAD(H5Awrite)
        DXG_H5_LOAD(H5Aread)
        DXG_H5_LOAD(H5Aget_space)
        DXG_H5_LOAD(H5Aget_type)
        DXG_H5_LOAD(H5Aclose)
#undef DXG_H5_LOAD
        // optional: direct chunk writes (1.10.3+), vlen reclaim (renamed in 1.12)
        api.H5Dwrite_chunk = reinterpret_cast<decltype(api.H5Dwrite_chunk)>(dlsym(api.handle, "H5Dwrite_chunk"));
        api.H5Dvlen_reclaim = reinterpret_cast<decltype(api.H5Dvlen_reclaim)>(dlsym(api.handle, "H5Dvlen_reclaim"));
        if (!api.H5Dvlen_reclaim)   // same arguments
            api.H5Dvlen_reclaim = reinterpret_cast<decltype(api.H5Dvlen_reclaim)>(dlsym(api.handle, "H5Treclaim"));
        for (const char* n : { "H5Literate2", "H5Literate", "H5Literate1" })
            if (!api.H5Literate)
                api.H5Literate = reinterpret_cast<decltype(api.H5Literate)>(dlsym(api.handle, n));
        if (!ok)
            return;
        if (!api.H5Literate || !api.H5Dvlen_reclaim) {
            load_err = "HDF5 library lacks H5Literate or H5Dvlen_reclaim";
            return;
        }
        if (api.H5open() < 0) {
            load_err = "H5open failed";
            return;
        }
        auto global = [&](const char* name, hid_t& out) {
            void* p = dlsym(api.handle, name);
            if (!p) {
                ok = false;
                load_err = std::string("HDF5 library lacks ") + name;
                return;
            }
            out = *static_cast<hid_t*>(p);
        };
        global("H5T_NATIVE_DOUBLE_g", api.f64);
        global("H5T_NATIVE_UINT64_g", api.u64);
        global("H5T_NATIVE_UINT8_g", api.u8);
        global("H5T_NATIVE_FLOAT_g", api.f32);
        global("H5T_NATIVE_INT32_g", api.i32);
        global("H5T_NATIVE_INT64_g", api.i64);
        global("H5T_C_S1_g", api.cs1);
        global("H5P_CLS_DATASET_CREATE_ID_g", api.dcreate);
        if (!ok)
            return;
        api.H5Eset_auto2(H5E_DEFAULT, nullptr, nullptr);   // errors come back as codes, not printed stacks
    });
    if (!load_err.empty()) {
        err = load_err;
        return nullptr;
    }
    return &api;
}

// Closes an HDF5 identifier when it goes out of scope.
struct Id {
    hid_t id = -1;
    herr_t (*closer)(hid_t) = nullptr;
    Id() = default;
    Id(hid_t i, herr_t (*c)(hid_t)) : id(i), closer(c) {}
    Id(const Id&) = delete;
    Id& operator=(const Id&) = delete;
    ~Id()
    {
        if (id >= 0 && closer)
            closer(id);
    }
    bool ok() const { return id >= 0; }
};

hid_t memType(const Api& a, int type)
{
    switch (type) {
    case DXG_H5_F64: return a.f64;
    case DXG_H5_U64: return a.u64;
    case DXG_H5_U8: return a.u8;
    case DXG_H5_F32: return a.f32;
    case DXG_H5_I32: return a.i32;
    case DXG_H5_I64: return a.i64;
    default: return -1;
    }
}

size_t typeSize(int type)
{
    switch (type) {
    case DXG_H5_F64:
    case DXG_H5_U64:
    case DXG_H5_I64: return 8;
    case DXG_H5_F32:
    case DXG_H5_I32: return 4;
    case DXG_H5_U8: return 1;
    default: return 0;
    }
}

// "/a/b" from "a/b", "/a//b/" ...; the components in `parts`.
std::string normalize(const char* path, std::vector<std::string>* parts = nullptr)
{
    std::string out;
    std::string cur;
    std::vector<std::string> p;
    for (const char* c = path ? path : ""; ; ++c) {
        if (*c == '/' || *c == 0) {
            if (!cur.empty()) {
                p.push_back(cur);
                out += "/" + cur;
                cur.clear();
            }
            if (*c == 0)
                break;
        } else {
            cur += *c;
        }
    }
    if (parts)
        *parts = p;
    return out.empty() ? "/" : out;
}

// ---------------------------------------------------------------- deflate

thread_local double g_deflate_ms = 0.0;
thread_local int g_deflate_threads = 0;

int zlibLevelFlags(int level)
{
    if (level < 2)
        return 0;
    if (level < 6)
        return 1;
    return level == 6 ? 2 : 3;
}

// One zlib stream (RFC 1950) of src, deflated by up to 16 threads.
bool parallelDeflate(const uint8_t* src, size_t n, int level, std::vector<uint8_t>& out)
{
    constexpr size_t kBlock = size_t(1) << 20;
    constexpr size_t kDict = 32768;
    const size_t nb = std::max<size_t>(1, (n + kBlock - 1) / kBlock);
    std::vector<std::vector<uint8_t>> parts(nb);
    std::vector<uLong> sums(nb);
    std::atomic<size_t> next { 0 };
    std::atomic<bool> failed { false };
    auto work = [&] {
        for (size_t b; (b = next.fetch_add(1)) < nb;) {
            const size_t off = b * kBlock, len = std::min(kBlock, n - std::min(n, off));
            z_stream zs {};
            if (deflateInit2(&zs, level, Z_DEFLATED, -15, 8, Z_DEFAULT_STRATEGY) != Z_OK) {
                failed = true;
                return;
            }
            if (b > 0) {
                const size_t d = std::min(kDict, off);
                deflateSetDictionary(&zs, src + off - d, static_cast<uInt>(d));
            }
            std::vector<uint8_t>& p = parts[b];
            p.resize(deflateBound(&zs, len) + 64);
            zs.next_in = const_cast<Bytef*>(src + off);
            zs.avail_in = static_cast<uInt>(len);
            zs.next_out = p.data();
            zs.avail_out = static_cast<uInt>(p.size());
            const bool last = b + 1 == nb;
            const int rc = deflate(&zs, last ? Z_FINISH : Z_SYNC_FLUSH);
            if ((last && rc != Z_STREAM_END) || (!last && (rc != Z_OK || zs.avail_in != 0)))
                failed = true;
            p.resize(p.size() - zs.avail_out);
            deflateEnd(&zs);
            sums[b] = adler32(adler32(0, nullptr, 0), src + off, static_cast<uInt>(len));
        }
    };
    const int threads = static_cast<int>(std::min<size_t>({ nb, 16, std::max(1u, std::thread::hardware_concurrency()) }));
    std::vector<std::thread> pool;
    for (int t = 1; t < threads; ++t)
        pool.emplace_back(work);
    work();
    for (auto& t : pool)
        t.join();
    g_deflate_threads = threads;
    if (failed)
        return false;
    uLong sum = adler32(0, nullptr, 0);
    size_t total = 6;
    for (size_t b = 0; b < nb; ++b) {
        const size_t len = std::min(kBlock, n - std::min(n, b * kBlock));
        sum = adler32_combine(sum, sums[b], static_cast<z_off_t>(len));
        total += parts[b].size();
    }
    out.clear();
    out.reserve(total);
    const int cmf = 0x78;   // deflate, 32 KiB window
    int flg = zlibLevelFlags(level) << 6;
    flg += 31 - (cmf * 256 + flg) % 31;
    out.push_back(static_cast<uint8_t>(cmf));
    out.push_back(static_cast<uint8_t>(flg));
    for (auto& p : parts)
        out.insert(out.end(), p.begin(), p.end());
    for (int s = 24; s >= 0; s -= 8)
        out.push_back(static_cast<uint8_t>((sum >> s) & 0xff));
    return true;
}

}  // namespace

struct dxg_h5 {
    const Api* api = nullptr;
    hid_t file = -1;
    bool writable = false;
};

namespace {

#define H5_API(f)                                                    \
    std::string load_err_;                                           \
    const Api* api_ = loadApi(load_err_);                            \
    if (!api_)                                                       \
        return fail(DXG_ERR_STATE, load_err_);                       \
    if (!(f) || (f)->file < 0)                                       \
        return fail(DXG_ERR_ARGUMENT, "HDF5 file is not open");      \
    const Api& A = *api_;                                            \
    std::lock_guard<std::mutex> lock_(g_mutex)

// 0 nothing, 1 group, 2 dataset, 3 other (caller holds the lock).
int kindOf(const Api& A, hid_t file, const std::string& path)
{
    if (path == "/")
        return 1;
    std::vector<std::string> parts;
    normalize(path.c_str(), &parts);
    std::string prefix;
    for (const auto& p : parts) {
        prefix += "/" + p;
        if (A.H5Lexists(file, prefix.c_str(), H5P_DEFAULT) <= 0)
            return 0;
    }
    Id o(A.H5Oopen(file, path.c_str(), H5P_DEFAULT), A.H5Oclose);
    if (!o.ok())
        return 0;
    const H5I_type_t t = A.H5Iget_type(o.id);
    return t == H5I_GROUP ? 1 : t == H5I_DATASET ? 2 : 3;
}

int makeGroups(const Api& A, hid_t file, const std::vector<std::string>& parts, size_t count)
{
    std::string prefix;
    for (size_t i = 0; i < count; ++i) {
        prefix += "/" + parts[i];
        const int k = kindOf(A, file, prefix);
        if (k == 1)
            continue;
        if (k != 0)
            return fail(DXG_ERR_ARGUMENT, "HDF5: " + prefix + " exists and is not a group");
        Id g(A.H5Gcreate2(file, prefix.c_str(), H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT), A.H5Gclose);
        if (!g.ok())
            return fail(DXG_ERR_DATA, "HDF5: cannot create group " + prefix);
    }
    return DXG_OK;
}

int prepareDataset(const Api& A, dxg_h5* f, const char* path, std::string& full)
{
    if (!f->writable)
        return fail(DXG_ERR_STATE, "HDF5 file is open read-only");
    std::vector<std::string> parts;
    full = normalize(path, &parts);
    if (parts.empty())
        return fail(DXG_ERR_ARGUMENT, "HDF5: empty dataset path");
    if (int rc = makeGroups(A, f->file, parts, parts.size() - 1); rc != DXG_OK)
        return rc;
    if (kindOf(A, f->file, full) != 0)
        return fail(DXG_ERR_ARGUMENT, "HDF5: " + full + " already exists");
    return DXG_OK;
}

herr_t collectName(hid_t, const char* name, const void*, void* data)
{
    static_cast<std::vector<std::string>*>(data)->emplace_back(name);
    return 0;
}

int packStrings(const std::vector<std::string>& v, char* buf, uint64_t cap, uint64_t* count, uint64_t* bytes)
{
    uint64_t need = 0;
    for (const auto& s : v)
        need += s.size() + 1;
    if (count)
        *count = v.size();
    if (bytes)
        *bytes = need;
    if (!buf)
        return DXG_OK;
    if (cap < need)
        return fail(DXG_ERR_ARGUMENT, "HDF5: string buffer too small: need " + std::to_string(need));
    for (const auto& s : v) {
        std::memcpy(buf, s.c_str(), s.size() + 1);
        buf += s.size() + 1;
    }
    return DXG_OK;
}

}  // namespace

extern "C" {

int dxg_h5_create(const char* path, dxg_h5** out)
{
    std::string err;
    const Api* a = loadApi(err);
    if (!a)
        return fail(DXG_ERR_STATE, err);
    if (!path || !out)
        return fail(DXG_ERR_ARGUMENT, "dxg_h5_create: null argument");
    std::lock_guard<std::mutex> lock(g_mutex);
    const hid_t fid = a->H5Fcreate(path, kAccTrunc, H5P_DEFAULT, H5P_DEFAULT);
    if (fid < 0)
        return fail(DXG_ERR_DATA, std::string("HDF5: cannot create ") + path);
    *out = new dxg_h5 { a, fid, true };
    return DXG_OK;
}

int dxg_h5_open(const char* path, dxg_h5** out)
{
    std::string err;
    const Api* a = loadApi(err);
    if (!a)
        return fail(DXG_ERR_STATE, err);
    if (!path || !out)
        return fail(DXG_ERR_ARGUMENT, "dxg_h5_open: null argument");
    std::lock_guard<std::mutex> lock(g_mutex);
    const hid_t fid = a->H5Fopen(path, kAccRdonly, H5P_DEFAULT);
    if (fid < 0)
        return fail(DXG_ERR_DATA, std::string("HDF5: cannot open ") + path);
    *out = new dxg_h5 { a, fid, false };
    return DXG_OK;
}

int dxg_h5_close(dxg_h5* f)
{
    if (!f)
        return DXG_OK;
    int rc = DXG_OK;
    if (f->file >= 0) {
        std::lock_guard<std::mutex> lock(g_mutex);
        if (f->api->H5Fclose(f->file) < 0)
            rc = fail(DXG_ERR_DATA, "HDF5: close failed");
    }
    delete f;
    return rc;
}

int dxg_h5_create_group(dxg_h5* f, const char* path)
{
    H5_API(f);
    if (!f->writable)
        return fail(DXG_ERR_STATE, "HDF5 file is open read-only");
    std::vector<std::string> parts;
    normalize(path, &parts);
    return makeGroups(A, f->file, parts, parts.size());
}

int dxg_h5_kind(dxg_h5* f, const char* path)
{
    H5_API(f);
    const int k = kindOf(A, f->file, normalize(path));
    return k == 3 ? 0 : k;
}

int dxg_h5_write_array(dxg_h5* f, const char* path, int type, int rank, const uint64_t* dims, const void* data, int deflate_level)
{
    H5_API(f);
    g_deflate_ms = 0.0;
    g_deflate_threads = 0;
    const hid_t mt = memType(A, type);
    if (mt < 0 || rank < 1 || rank > 32 || !dims || deflate_level > 9)
        return fail(DXG_ERR_ARGUMENT, "dxg_h5_write_array: bad type, rank or deflate level");
    uint64_t count = 1;
    for (int i = 0; i < rank; ++i)
        count *= dims[i];
    if (count > 0 && !data)
        return fail(DXG_ERR_ARGUMENT, "dxg_h5_write_array: null data");
    std::string full;
    if (int rc = prepareDataset(A, f, path, full); rc != DXG_OK)
        return rc;
    std::vector<hsize_t> hd(dims, dims + rank);
    Id space(A.H5Screate_simple(rank, hd.data(), nullptr), A.H5Sclose);
    if (!space.ok())
        return fail(DXG_ERR_ARGUMENT, "HDF5: bad extent for " + full);
    const size_t bytes = count * typeSize(type);
    if (deflate_level < 0) {
        Id d(A.H5Dcreate2(f->file, full.c_str(), mt, space.id, H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT), A.H5Dclose);
        if (!d.ok() || A.H5Dwrite(d.id, mt, H5S_ALL, H5S_ALL, H5P_DEFAULT, data) < 0)
            return fail(DXG_ERR_DATA, "HDF5: cannot write " + full);
        return DXG_OK;
    }
    if (count == 0)
        return fail(DXG_ERR_ARGUMENT, "HDF5: a compressed dataset needs at least one element");
    if (bytes >= (uint64_t(1) << 32))
        return fail(DXG_ERR_ARGUMENT, "HDF5: one chunk holds less than 4 GiB; " + full + " has " + std::to_string(bytes) + " bytes");
    Id plist(A.H5Pcreate(A.dcreate), A.H5Pclose);
    if (!plist.ok() || A.H5Pset_chunk(plist.id, rank, hd.data()) < 0 || A.H5Pset_deflate(plist.id, static_cast<unsigned>(deflate_level)) < 0)
        return fail(DXG_ERR_DATA, "HDF5: cannot set chunking / deflate for " + full);
    Id d(A.H5Dcreate2(f->file, full.c_str(), mt, space.id, H5P_DEFAULT, plist.id, H5P_DEFAULT), A.H5Dclose);
    if (!d.ok())
        return fail(DXG_ERR_DATA, "HDF5: cannot create " + full);
    // the library's own (single-threaded) filter: older libraries, or on
    // request (DXG_H5_LIBRARY_DEFLATE=1) to compare with the reference's path
    const char* lib_env = std::getenv("DXG_H5_LIBRARY_DEFLATE");
    const bool library_deflate = lib_env && lib_env[0] == '1';
    if (!A.H5Dwrite_chunk || library_deflate) {
        const auto t0 = std::chrono::steady_clock::now();
        if (A.H5Dwrite(d.id, mt, H5S_ALL, H5S_ALL, H5P_DEFAULT, data) < 0)
            return fail(DXG_ERR_DATA, "HDF5: cannot write " + full);
        g_deflate_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
        g_deflate_threads = 1;
        return DXG_OK;
    }
    const auto t0 = std::chrono::steady_clock::now();
    std::vector<uint8_t> z;
    if (!parallelDeflate(static_cast<const uint8_t*>(data), bytes, deflate_level, z))
        return fail(DXG_ERR_DATA, "HDF5: deflate failed for " + full);
    g_deflate_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    std::vector<hsize_t> origin(rank, 0);
    if (A.H5Dwrite_chunk(d.id, H5P_DEFAULT, 0u, origin.data(), z.size(), z.data()) < 0)
        return fail(DXG_ERR_DATA, "HDF5: cannot write the chunk of " + full);
    return DXG_OK;
}

int dxg_h5_write_strings(dxg_h5* f, const char* path, const char* const* strings, uint64_t n)
{
    H5_API(f);
    if (n > 0 && !strings)
        return fail(DXG_ERR_ARGUMENT, "dxg_h5_write_strings: null strings");
    std::string full;
    if (int rc = prepareDataset(A, f, path, full); rc != DXG_OK)
        return rc;
    Id t(A.H5Tcopy(A.cs1), A.H5Tclose);
    if (!t.ok() || A.H5Tset_size(t.id, H5T_VARIABLE) < 0)
        return fail(DXG_ERR_DATA, "HDF5: cannot make a variable-length string type");
    hsize_t dim = n;
    Id space(A.H5Screate_simple(1, &dim, nullptr), A.H5Sclose);
    Id d(A.H5Dcreate2(f->file, full.c_str(), t.id, space.id, H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT), A.H5Dclose);
    if (!d.ok())
        return fail(DXG_ERR_DATA, "HDF5: cannot create " + full);
    std::vector<const char*> p(strings, strings + n);
    for (auto& s : p)
        if (!s)
            s = "";
    if (n > 0 && A.H5Dwrite(d.id, t.id, H5S_ALL, H5S_ALL, H5P_DEFAULT, p.data()) < 0)
        return fail(DXG_ERR_DATA, "HDF5: cannot write " + full);
    return DXG_OK;
}

int dxg_h5_write_attribute(dxg_h5* f, const char* path, const char* name, int type, const void* data, uint64_t n, int scalar)
{
    H5_API(f);
    if (!f->writable)
        return fail(DXG_ERR_STATE, "HDF5 file is open read-only");
    if (!name || !data || (type != DXG_H5_F64 && type != DXG_H5_U64) || n == 0 || (scalar && n != 1))
        return fail(DXG_ERR_ARGUMENT, "dxg_h5_write_attribute: bad name, data, type or count");
    std::vector<std::string> parts;
    const std::string full = normalize(path, &parts);
    if (int rc = makeGroups(A, f->file, parts, parts.size()); rc != DXG_OK)
        return rc;
    Id g(A.H5Gopen2(f->file, full.c_str(), H5P_DEFAULT), A.H5Gclose);
    if (!g.ok())
        return fail(DXG_ERR_DATA, "HDF5: cannot open group " + full);
    hsize_t dim = n;
    Id space(scalar ? A.H5Screate(H5S_SCALAR) : A.H5Screate_simple(1, &dim, nullptr), A.H5Sclose);
    const hid_t mt = memType(A, type);
    Id at(A.H5Acreate2(g.id, name, mt, space.id, H5P_DEFAULT, H5P_DEFAULT), A.H5Aclose);
    if (!at.ok() || A.H5Awrite(at.id, mt, data) < 0)
        return fail(DXG_ERR_DATA, "HDF5: cannot write attribute " + std::string(name) + " of " + full);
    return DXG_OK;
}

int dxg_h5_dataset_shape(dxg_h5* f, const char* path, int* rank, uint64_t* dims, int* type_class, uint64_t* type_size)
{
    H5_API(f);
    const std::string full = normalize(path);
    if (kindOf(A, f->file, full) != 2)
        return fail(DXG_ERR_ARGUMENT, "HDF5: no dataset at " + full);
    Id d(A.H5Dopen2(f->file, full.c_str(), H5P_DEFAULT), A.H5Dclose);
    Id space(A.H5Dget_space(d.id), A.H5Sclose);
    Id t(A.H5Dget_type(d.id), A.H5Tclose);
    const int r = A.H5Sget_simple_extent_ndims(space.id);
    if (r < 0 || r > 32)
        return fail(DXG_ERR_DATA, "HDF5: unsupported extent for " + full);
    hsize_t hd[32];
    A.H5Sget_simple_extent_dims(space.id, hd, nullptr);
    if (rank)
        *rank = r;
    if (dims)
        for (int i = 0; i < r; ++i)
            dims[i] = hd[i];
    if (type_class)
        *type_class = static_cast<int>(A.H5Tget_class(t.id));
    if (type_size)
        *type_size = A.H5Tis_variable_str(t.id) > 0 ? 0 : A.H5Tget_size(t.id);
    return DXG_OK;
}

int dxg_h5_read_array(dxg_h5* f, const char* path, int type, void* out, uint64_t n)
{
    H5_API(f);
    const hid_t mt = memType(A, type);
    if (mt < 0)
        return fail(DXG_ERR_ARGUMENT, "dxg_h5_read_array: bad type");
    const std::string full = normalize(path);
    if (kindOf(A, f->file, full) != 2)
        return fail(DXG_ERR_ARGUMENT, "HDF5: no dataset at " + full);
    Id d(A.H5Dopen2(f->file, full.c_str(), H5P_DEFAULT), A.H5Dclose);
    Id space(A.H5Dget_space(d.id), A.H5Sclose);
    const hssize_t count = A.H5Sget_simple_extent_npoints(space.id);
    if (count < 0 || static_cast<uint64_t>(count) != n)
        return fail(DXG_ERR_ARGUMENT, "HDF5: " + full + " holds " + std::to_string(count) + " elements, not " + std::to_string(n));
    if (n == 0)
        return DXG_OK;
    if (!out)
        return fail(DXG_ERR_ARGUMENT, "dxg_h5_read_array: null output");
    if (A.H5Dread(d.id, mt, H5S_ALL, H5S_ALL, H5P_DEFAULT, out) < 0)
        return fail(DXG_ERR_DATA, "HDF5: cannot read " + full + " as the requested type");
    return DXG_OK;
}

int dxg_h5_read_strings(dxg_h5* f, const char* path, char* buf, uint64_t cap, uint64_t* count, uint64_t* bytes)
{
    H5_API(f);
    const std::string full = normalize(path);
    if (kindOf(A, f->file, full) != 2)
        return fail(DXG_ERR_ARGUMENT, "HDF5: no dataset at " + full);
    Id d(A.H5Dopen2(f->file, full.c_str(), H5P_DEFAULT), A.H5Dclose);
    Id space(A.H5Dget_space(d.id), A.H5Sclose);
    Id ft(A.H5Dget_type(d.id), A.H5Tclose);
    if (A.H5Tget_class(ft.id) != H5T_STRING)
        return fail(DXG_ERR_ARGUMENT, "HDF5: " + full + " does not hold strings");
    const hssize_t n = A.H5Sget_simple_extent_npoints(space.id);
    std::vector<std::string> v;
    if (n > 0 && A.H5Tis_variable_str(ft.id) > 0) {
        Id mt(A.H5Tcopy(A.cs1), A.H5Tclose);
        A.H5Tset_size(mt.id, H5T_VARIABLE);
        std::vector<char*> p(static_cast<size_t>(n), nullptr);
        if (A.H5Dread(d.id, mt.id, H5S_ALL, H5S_ALL, H5P_DEFAULT, p.data()) < 0)
            return fail(DXG_ERR_DATA, "HDF5: cannot read " + full);
        for (char* s : p)
            v.emplace_back(s ? s : "");
        A.H5Dvlen_reclaim(mt.id, space.id, H5P_DEFAULT, p.data());
    } else if (n > 0) {   // fixed-length strings
        const size_t sz = A.H5Tget_size(ft.id);
        Id mt(A.H5Tcopy(A.cs1), A.H5Tclose);
        A.H5Tset_size(mt.id, sz);
        std::vector<char> raw(sz * static_cast<size_t>(n));
        if (A.H5Dread(d.id, mt.id, H5S_ALL, H5S_ALL, H5P_DEFAULT, raw.data()) < 0)
            return fail(DXG_ERR_DATA, "HDF5: cannot read " + full);
        for (hssize_t i = 0; i < n; ++i) {
            const char* s = raw.data() + i * sz;
            v.emplace_back(s, strnlen(s, sz));
        }
    }
    return packStrings(v, buf, cap, count, bytes);
}

int dxg_h5_read_attribute(dxg_h5* f, const char* path, const char* name, int type, void* out, uint64_t cap, uint64_t* n, int* rank)
{
    H5_API(f);
    if (!name || !n || (type != DXG_H5_F64 && type != DXG_H5_U64))
        return fail(DXG_ERR_ARGUMENT, "dxg_h5_read_attribute: bad name, type or count pointer");
    *n = 0;
    if (rank)
        *rank = 0;
    const std::string full = normalize(path);
    const int k = kindOf(A, f->file, full);
    if (k != 1 && k != 2)
        return DXG_OK;
    Id o(A.H5Oopen(f->file, full.c_str(), H5P_DEFAULT), A.H5Oclose);
    if (A.H5Aexists(o.id, name) <= 0)
        return DXG_OK;
    Id at(A.H5Aopen(o.id, name, H5P_DEFAULT), A.H5Aclose);
    Id space(A.H5Aget_space(at.id), A.H5Sclose);
    Id t(A.H5Aget_type(at.id), A.H5Tclose);
    const int r = A.H5Sget_simple_extent_ndims(space.id);
    const hssize_t count = A.H5Sget_simple_extent_npoints(space.id);
    const H5T_class_t cls = A.H5Tget_class(t.id);
    const size_t sz = A.H5Tget_size(t.id);
    // loadAttribute: rank <= 1 and an 8-byte float / integer, else absent
    if (r > 1 || count < 1 || sz != 8 || cls != (type == DXG_H5_F64 ? H5T_FLOAT : H5T_INTEGER))
        return DXG_OK;
    if (rank)
        *rank = r;
    if (!out) {
        *n = static_cast<uint64_t>(count);
        return DXG_OK;
    }
    if (cap < static_cast<uint64_t>(count))
        return fail(DXG_ERR_ARGUMENT, "HDF5: attribute " + std::string(name) + " has " + std::to_string(count) + " values");
    if (A.H5Aread(at.id, memType(A, type), out) < 0)
        return fail(DXG_ERR_DATA, "HDF5: cannot read attribute " + std::string(name));
    *n = static_cast<uint64_t>(count);
    return DXG_OK;
}

int dxg_h5_list(dxg_h5* f, const char* path, char* buf, uint64_t cap, uint64_t* count, uint64_t* bytes)
{
    H5_API(f);
    const std::string full = normalize(path);
    if (kindOf(A, f->file, full) != 1)
        return fail(DXG_ERR_ARGUMENT, "HDF5: no group at " + full);
    Id g(A.H5Gopen2(f->file, full.c_str(), H5P_DEFAULT), A.H5Gclose);
    std::vector<std::string> names;
    hsize_t idx = 0;
    if (A.H5Literate(g.id, H5_INDEX_NAME, H5_ITER_INC, &idx, reinterpret_cast<H5L_iterate_t>(collectName), &names) < 0)
        return fail(DXG_ERR_DATA, "HDF5: cannot list " + full);
    return packStrings(names, buf, cap, count, bytes);
}

double dxg_h5_last_deflate_ms(int* threads)
{
    if (threads)
        *threads = g_deflate_threads;
    return g_deflate_ms;
}

}  // extern "C"
