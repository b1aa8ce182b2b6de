// Read-only NetCDF access for the grid-file readers (src/Grids/FileGrids.jl uses
// NCDatasets.jl, i.e. libnetcdf, for the same purpose).
//
// * Classic formats (CDF-1 classic, CDF-2 64-bit offset, CDF-5 64-bit data) are parsed
//   here from the published format specification: a big-endian header (dimensions,
//   global attributes, variables with their offsets) followed by the data of the
//   fixed-size variables and the interleaved records of the record variables.
// * NetCDF-4 files are HDF5 files: variables are the datasets of the root group (minus
//   the pure dimension scales netCDF-4 writes) and global attributes are the root
//   group's attributes.  They are read through the HDF5 C library, loaded with dlopen
//   on first use (DMPLEX_HDF5_LIBRARY, else the system's libhdf5), so the library
//   builds and runs without HDF5 for classic files.
#include <dlfcn.h>
#include <fcntl.h>
#include <sys/stat.h>
#include <unistd.h>

#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <string>
#include <type_traits>
#include <vector>

#include "common.hpp"

namespace dmx {
namespace {

enum NcType { NC_BYTE = 1, NC_CHAR, NC_SHORT, NC_INT, NC_FLOAT, NC_DOUBLE, NC_UBYTE, NC_USHORT, NC_UINT,
              NC_INT64, NC_UINT64 };

int type_size(int t) {
    switch (t) {
        case NC_BYTE: case NC_CHAR: case NC_UBYTE: return 1;
        case NC_SHORT: case NC_USHORT: return 2;
        case NC_INT: case NC_UINT: case NC_FLOAT: return 4;
        case NC_DOUBLE: case NC_INT64: case NC_UINT64: return 8;
        default: return 0;
    }
}

bool is_integer(int t) { return t != NC_CHAR && t != NC_FLOAT && t != NC_DOUBLE; }

// Big-endian value i of a raw buffer of type t, as double / int64.
template <class T>
T be_value(const uint8_t* p, int t) {
    uint64_t u = 0;
    const int n = type_size(t);
    for (int k = 0; k < n; ++k) u = (u << 8) | p[k];
    switch (t) {
        case NC_BYTE: return T(int8_t(u));
        case NC_CHAR: case NC_UBYTE: return T(uint8_t(u));
        case NC_SHORT: return T(int16_t(u));
        case NC_USHORT: return T(uint16_t(u));
        case NC_INT: return T(int32_t(u));
        case NC_UINT: return T(uint32_t(u));
        case NC_INT64: return T(int64_t(u));
        case NC_UINT64: return T(u);
        case NC_FLOAT: { uint32_t w = uint32_t(u); float f; std::memcpy(&f, &w, 4); return T(f); }
        default: { double d; std::memcpy(&d, &u, 8); return T(d); }
    }
}

struct Var {
    std::string name;
    int type = 0;
    std::vector<int64_t> shape;  // C order, as ncdump prints it
    bool record = false;         // first dimension is the record (unlimited) dimension
    int64_t vsize = 0;           // bytes per record (record variables) or in total
    int64_t begin = 0;
};

struct Attr {
    std::string name;
    int type = 0;
    std::vector<uint8_t> raw;  // big-endian values
};

// --- HDF5 through dlopen ---------------------------------------------------------------
using hid_t = int64_t;
using herr_t = int;
using htri_t = int;
using hsize_t = unsigned long long;
struct H5GInfo {
    int storage_type;
    hsize_t nlinks;
    int64_t max_corder;
    bool mounted;
};
constexpr hid_t kH5P_DEFAULT = 0, kH5S_ALL = 0, kH5E_DEFAULT = 0;
constexpr int kH5T_INTEGER = 0, kH5T_FLOAT = 1, kH5T_STRING = 3, kH5T_SGN_NONE = 0;

struct Hdf5 {
    void* so = nullptr;
    herr_t (*open)();
    hid_t (*Fopen)(const char*, unsigned, hid_t);
    herr_t (*Fclose)(hid_t);
    herr_t (*Eset_auto2)(hid_t, void*, void*);
    herr_t (*Gget_info)(hid_t, H5GInfo*);
    long (*Lget_name_by_idx)(hid_t, const char*, int, int, hsize_t, char*, size_t, hid_t);
    htri_t (*Lexists)(hid_t, const char*, hid_t);
    hid_t (*Dopen2)(hid_t, const char*, hid_t);
    herr_t (*Dclose)(hid_t);
    hid_t (*Dget_space)(hid_t);
    hid_t (*Dget_type)(hid_t);
    herr_t (*Dread)(hid_t, hid_t, hid_t, hid_t, hid_t, void*);
    int (*Sget_simple_extent_ndims)(hid_t);
    int (*Sget_simple_extent_dims)(hid_t, hsize_t*, hsize_t*);
    herr_t (*Sclose)(hid_t);
    int (*Tget_class)(hid_t);
    size_t (*Tget_size)(hid_t);
    int (*Tget_sign)(hid_t);
    htri_t (*Tis_variable_str)(hid_t);
    hid_t (*Tcopy)(hid_t);
    herr_t (*Tset_size)(hid_t, size_t);
    herr_t (*Tclose)(hid_t);
    htri_t (*Aexists)(hid_t, const char*);
    hid_t (*Aopen)(hid_t, const char*, hid_t);
    herr_t (*Aclose)(hid_t);
    hid_t (*Aget_type)(hid_t);
    hid_t (*Aget_space)(hid_t);
    herr_t (*Aread)(hid_t, hid_t, void*);
    herr_t (*free_memory)(void*);
    hid_t native_double = -1, native_llong = -1, c_s1 = -1;
};

// Why the HDF5 library could not be used (empty while it can).
std::string hdf5_error;

Hdf5* hdf5() {
    static Hdf5 h;
    static bool tried = false;
    if (tried) return h.so ? &h : nullptr;
    tried = true;
    std::vector<std::string> names;
    if (const char* env = std::getenv("DMPLEX_HDF5_LIBRARY")) names.push_back(env);
    for (const char* n : {"libhdf5.so", "libhdf5_serial.so", "libhdf5.so.310", "libhdf5.so.200", "libhdf5.so.103",
                          "libhdf5_serial.so.103", "libhdf5.so.101", "/opt/conda/lib/libhdf5.so"})
        names.push_back(n);
    for (const auto& n : names)
        if ((h.so = dlopen(n.c_str(), RTLD_NOW | RTLD_LOCAL))) break;
    if (!h.so) {
        hdf5_error = "the HDF5 library (libhdf5.so) was not found; set DMPLEX_HDF5_LIBRARY to its path";
        return nullptr;
    }
    bool ok = true;
    auto sym = [&](auto& fn, const char* name) {
        void* p = dlsym(h.so, name);
        if (!p) ok = false;
        fn = reinterpret_cast<std::remove_reference_t<decltype(fn)>>(p);
    };
    sym(h.open, "H5open");
    sym(h.Fopen, "H5Fopen");
    sym(h.Fclose, "H5Fclose");
    sym(h.Eset_auto2, "H5Eset_auto2");
    sym(h.Gget_info, "H5Gget_info");
    sym(h.Lget_name_by_idx, "H5Lget_name_by_idx");
    sym(h.Lexists, "H5Lexists");
    sym(h.Dopen2, "H5Dopen2");
    sym(h.Dclose, "H5Dclose");
    sym(h.Dget_space, "H5Dget_space");
    sym(h.Dget_type, "H5Dget_type");
    sym(h.Dread, "H5Dread");
    sym(h.Sget_simple_extent_ndims, "H5Sget_simple_extent_ndims");
    sym(h.Sget_simple_extent_dims, "H5Sget_simple_extent_dims");
    sym(h.Sclose, "H5Sclose");
    sym(h.Tget_class, "H5Tget_class");
    sym(h.Tget_size, "H5Tget_size");
    sym(h.Tget_sign, "H5Tget_sign");
    sym(h.Tis_variable_str, "H5Tis_variable_str");
    sym(h.Tcopy, "H5Tcopy");
    sym(h.Tset_size, "H5Tset_size");
    sym(h.Tclose, "H5Tclose");
    sym(h.Aexists, "H5Aexists");
    sym(h.Aopen, "H5Aopen");
    sym(h.Aclose, "H5Aclose");
    sym(h.Aget_type, "H5Aget_type");
    sym(h.Aget_space, "H5Aget_space");
    sym(h.Aread, "H5Aread");
    sym(h.free_memory, "H5free_memory");
    const hid_t *nd = nullptr, *nl = nullptr, *cs = nullptr;
    nd = static_cast<const hid_t*>(dlsym(h.so, "H5T_NATIVE_DOUBLE_g"));
    nl = static_cast<const hid_t*>(dlsym(h.so, "H5T_NATIVE_LLONG_g"));
    cs = static_cast<const hid_t*>(dlsym(h.so, "H5T_C_S1_g"));
    if (!ok || !nd || !nl || !cs || h.open() < 0) {
        hdf5_error = "the HDF5 library lacks part of the 1.10+ C API";
        dlclose(h.so);
        h.so = nullptr;
        return nullptr;
    }
    // The type ids are set by H5open; error stacks are reported here, not printed.
    h.native_double = *nd;
    h.native_llong = *nl;
    h.c_s1 = *cs;
    h.Eset_auto2(kH5E_DEFAULT, nullptr, nullptr);
    return &h;
}

}  // namespace

struct NcFile {
    int format = 0;  // 1, 2, 5 (classic family) or 4 (NetCDF-4 / HDF5)
    std::string path;
    int fd = -1;
    int64_t file_size = 0;
    int64_t numrecs = 0;
    int64_t recsize = 0;
    std::vector<Var> vars;
    std::vector<Attr> gatts;
    hid_t h5 = -1;
    ~NcFile() {
        if (fd >= 0) ::close(fd);
        if (h5 >= 0) hdf5()->Fclose(h5);
    }
    int find(const char* name) const {
        for (size_t i = 0; i < vars.size(); ++i)
            if (vars[i].name == name) return int(i);
        return -1;
    }
};

namespace {

// Sequential big-endian reader over the header bytes.
struct Cursor {
    const uint8_t* p;
    int64_t n, pos = 0;
    bool ok = true;
    uint64_t be(int bytes) {
        if (pos + bytes > n) {
            ok = false;
            return 0;
        }
        uint64_t u = 0;
        for (int k = 0; k < bytes; ++k) u = (u << 8) | p[pos + k];
        pos += bytes;
        return u;
    }
    std::string text(int64_t len) {
        if (len < 0 || pos + len > n) {
            ok = false;
            return {};
        }
        std::string s(reinterpret_cast<const char*>(p + pos), size_t(len));
        pos += (len + 3) / 4 * 4;
        if (pos > n) ok = false;
        return s;
    }
};

bool pread_all(int fd, void* buf, int64_t n, int64_t off) {
    auto* b = static_cast<uint8_t*>(buf);
    while (n > 0) {
        ssize_t r = ::pread(fd, b, size_t(n), off);
        if (r <= 0) return false;
        b += r;
        n -= r;
        off += r;
    }
    return true;
}

// Parses the classic header (CDF-1/2/5).  Returns an error text, empty on success.
std::string parse_classic(NcFile& f, const std::vector<uint8_t>& head) {
    Cursor c{head.data(), int64_t(head.size())};
    c.pos = 4;
    const int big = f.format == 5 ? 8 : 4;          // NON_NEG counts, dimension ids, vsize
    const int off = f.format == 1 ? 4 : 8;          // variable offsets
    uint64_t nrec = c.be(big);
    const bool streaming = f.format == 5 ? nrec == ~0ull : nrec == 0xffffffffull;
    f.numrecs = streaming ? -1 : int64_t(nrec);
    // dimensions
    std::vector<std::string> dim_names;
    std::vector<int64_t> dim_len;
    int64_t record_dim = -1;
    uint64_t tag = c.be(4), count = c.be(big);
    if (tag != 0x0A && !(tag == 0 && count == 0)) return "bad dimension list";
    for (uint64_t i = 0; i < count && c.ok; ++i) {
        int64_t len = int64_t(c.be(big));
        dim_names.push_back(c.text(len));
        dim_len.push_back(int64_t(c.be(big)));
        if (dim_len.back() == 0) {
            if (record_dim >= 0) return "more than one record dimension";
            record_dim = int64_t(i);
        }
    }
    auto read_attrs = [&](std::vector<Attr>* out) -> bool {
        uint64_t t = c.be(4), m = c.be(big);
        if (t != 0x0C && !(t == 0 && m == 0)) return false;
        for (uint64_t i = 0; i < m && c.ok; ++i) {
            Attr a;
            a.name = c.text(int64_t(c.be(big)));
            a.type = int(c.be(4));
            int64_t nelem = int64_t(c.be(big));
            int ts = type_size(a.type);
            if (!ts || nelem < 0 || c.pos + nelem * ts > c.n) return false;
            a.raw.assign(c.p + c.pos, c.p + c.pos + nelem * ts);
            c.pos += (nelem * ts + 3) / 4 * 4;
            if (out) out->push_back(std::move(a));
        }
        return c.ok;
    };
    if (!read_attrs(&f.gatts)) return "bad global attribute list";
    tag = c.be(4);
    count = c.be(big);
    if (tag != 0x0B && !(tag == 0 && count == 0)) return "bad variable list";
    int nrecvars = 0;
    for (uint64_t i = 0; i < count && c.ok; ++i) {
        Var v;
        v.name = c.text(int64_t(c.be(big)));
        uint64_t nd = c.be(big);
        if (nd > 64) return "variable " + v.name + " has too many dimensions";
        for (uint64_t k = 0; k < nd && c.ok; ++k) {
            uint64_t id = c.be(big);
            if (id >= dim_len.size()) return "variable " + v.name + " names an unknown dimension";
            if (int64_t(id) == record_dim) {
                if (k != 0) return "variable " + v.name + ": the record dimension is not its first";
                v.record = true;
            }
            v.shape.push_back(dim_len[id]);
        }
        if (!read_attrs(nullptr)) return "bad attribute list of variable " + v.name;
        v.type = int(c.be(4));
        if (!type_size(v.type) || (f.format != 5 && v.type > NC_DOUBLE))
            return "variable " + v.name + " has an unknown type";
        v.vsize = int64_t(c.be(big));
        v.begin = int64_t(c.be(off));
        nrecvars += v.record;
        f.vars.push_back(std::move(v));
    }
    if (!c.ok) return "truncated header";
    // Record layout: one record holds every record variable's slab, each padded to 4 bytes
    // unless there is a single record variable.  The header's vsize is advisory (it
    // saturates for huge variables), so sizes are computed from the shapes.
    f.recsize = 0;
    for (auto& v : f.vars) {
        if (!v.record) continue;
        int64_t slab = type_size(v.type);
        for (size_t k = 1; k < v.shape.size(); ++k) slab *= v.shape[k];
        f.recsize += nrecvars == 1 ? slab : (slab + 3) / 4 * 4;
    }
    if (f.numrecs < 0) {  // streaming: as many whole records as the file holds
        int64_t first = -1;
        for (auto& v : f.vars)
            if (v.record && (first < 0 || v.begin < first)) first = v.begin;
        f.numrecs = (first < 0 || f.recsize == 0) ? 0 : std::max<int64_t>(0, (f.file_size - first) / f.recsize);
    }
    for (auto& v : f.vars)
        if (v.record) v.shape[0] = f.numrecs;
    for (auto& v : f.vars) {
        int64_t n = type_size(v.type);
        for (size_t k = v.record ? 1 : 0; k < v.shape.size(); ++k) n *= v.shape[k];
        if (v.record && f.numrecs == 0) continue;  // nothing to read; its records may start past the end
        int64_t end = v.begin + (v.record ? (f.numrecs - 1) * f.recsize + n : n);
        if (v.begin < 0 || end > f.file_size) return "variable " + v.name + " extends past the end of the file";
    }
    return {};
}

// Datasets of the root group that are netCDF variables (pure dimension scales, whose
// NAME attribute starts with "This is a netCDF dimension but not a netCDF variable",
// are not).
std::string read_string_attr(Hdf5* H, hid_t obj, const char* name, bool* found);

std::string parse_hdf5(NcFile& f) {
    Hdf5* H = hdf5();
    if (!H) return "NetCDF-4 (HDF5) file, and " + hdf5_error;
    f.h5 = H->Fopen(f.path.c_str(), 0u, kH5P_DEFAULT);
    if (f.h5 < 0) return "HDF5 could not open the file";
    H5GInfo info{};
    if (H->Gget_info(f.h5, &info) < 0) return "HDF5 could not list the root group";
    for (hsize_t i = 0; i < info.nlinks; ++i) {
        char name[1024];
        long len = H->Lget_name_by_idx(f.h5, ".", 0, 0, i, name, sizeof(name), kH5P_DEFAULT);
        if (len < 0 || len >= long(sizeof(name))) continue;
        hid_t d = H->Dopen2(f.h5, name, kH5P_DEFAULT);
        if (d < 0) continue;  // a group or a dangling link
        bool has_name = false;
        std::string tag = read_string_attr(H, d, "NAME", &has_name);
        const bool pure_dim = has_name && tag.rfind("This is a netCDF dimension but not a netCDF variable", 0) == 0;
        Var v;
        v.name = name;
        hid_t s = H->Dget_space(d), t = H->Dget_type(d);
        int nd = s >= 0 ? H->Sget_simple_extent_ndims(s) : -1;
        std::vector<hsize_t> dims(size_t(std::max(nd, 0)) + 1);
        if (nd >= 0 && H->Sget_simple_extent_dims(s, dims.data(), nullptr) >= 0)
            for (int k = 0; k < nd; ++k) v.shape.push_back(int64_t(dims[k]));
        const int cls = t >= 0 ? H->Tget_class(t) : -1;
        const size_t sz = t >= 0 ? H->Tget_size(t) : 0;
        if (cls == kH5T_FLOAT) {
            v.type = sz == 4 ? NC_FLOAT : NC_DOUBLE;
        } else if (cls == kH5T_INTEGER) {
            const bool u = H->Tget_sign(t) == kH5T_SGN_NONE;
            v.type = sz == 1 ? (u ? NC_UBYTE : NC_BYTE) : sz == 2 ? (u ? NC_USHORT : NC_SHORT)
                   : sz == 4 ? (u ? NC_UINT : NC_INT) : (u ? NC_UINT64 : NC_INT64);
        } else if (cls == kH5T_STRING) {
            v.type = NC_CHAR;
        }
        if (t >= 0) H->Tclose(t);
        if (s >= 0) H->Sclose(s);
        H->Dclose(d);
        if (nd >= 0 && v.type != 0 && !pure_dim) f.vars.push_back(std::move(v));
    }
    return {};
}

}  // namespace

namespace {

std::string read_string_attr(Hdf5* H, hid_t obj, const char* name, bool* found) {
    *found = false;
    if (H->Aexists(obj, name) <= 0) return {};
    hid_t a = H->Aopen(obj, name, kH5P_DEFAULT);
    if (a < 0) return {};
    hid_t t = H->Aget_type(a);
    std::string out;
    if (t >= 0 && H->Tget_class(t) == kH5T_STRING) {
        *found = true;
        hid_t m = H->Tcopy(H->c_s1);
        if (H->Tis_variable_str(t) > 0) {
            H->Tset_size(m, size_t(-1));  // H5T_VARIABLE
            char* p = nullptr;
            if (H->Aread(a, m, &p) >= 0 && p) {
                out = p;
                H->free_memory(p);
            }
        } else {
            const size_t n = H->Tget_size(t);
            H->Tset_size(m, n);
            std::vector<char> buf(n + 1, 0);
            if (H->Aread(a, m, buf.data()) >= 0) out.assign(buf.data(), strnlen(buf.data(), n));
        }
        H->Tclose(m);
    }
    if (t >= 0) H->Tclose(t);
    H->Aclose(a);
    return out;
}

int open_file(const char* path, NcFile** out) {
    *out = nullptr;
    auto f = std::make_unique<NcFile>();
    f->path = path;
    f->fd = ::open(path, O_RDONLY);
    if (f->fd < 0) return fail(-5, std::string(path) + ": no such file or not readable");
    struct stat st{};
    fstat(f->fd, &st);
    f->file_size = int64_t(st.st_size);
    uint8_t magic[8] = {0};
    if (f->file_size < 4 || !pread_all(f->fd, magic, std::min<int64_t>(8, f->file_size), 0))
        return fail(-2, std::string(path) + " is not a NetCDF file");
    static const uint8_t hdf5_sig[8] = {0x89, 'H', 'D', 'F', '\r', '\n', 0x1a, '\n'};
    if (f->file_size >= 8 && std::memcmp(magic, hdf5_sig, 8) == 0) {
        f->format = 4;
        ::close(f->fd);
        f->fd = -1;
        std::string err = parse_hdf5(*f);
        if (!err.empty()) return fail(-2, std::string(path) + ": " + err);
        *out = f.release();
        return 0;
    }
    if (magic[0] != 'C' || magic[1] != 'D' || magic[2] != 'F' || (magic[3] != 1 && magic[3] != 2 && magic[3] != 5))
        return fail(-2, std::string(path) + " is not a NetCDF file");
    f->format = magic[3];
    // The header ends where the first variable's data begins; read it in growing chunks.
    std::vector<uint8_t> head;
    std::string err;
    for (int64_t want = 1 << 16;; want *= 4) {
        want = std::min(want, f->file_size);
        head.resize(size_t(want));
        if (!pread_all(f->fd, head.data(), want, 0)) return fail(-2, std::string(path) + ": read error");
        f->vars.clear();
        f->gatts.clear();
        err = parse_classic(*f, head);
        if (err.empty() || err != "truncated header" || want == f->file_size) break;
    }
    if (!err.empty()) return fail(-2, std::string(path) + ": " + err);
    *out = f.release();
    return 0;
}

const Var* lookup(const NcFile* f, const char* name) {
    int i = f->find(name);
    if (i < 0) {
        fail(-3, "no variable " + std::string(name) + " in " + f->path);
        return nullptr;
    }
    return &f->vars[size_t(i)];
}

int64_t count_of(const Var& v) {
    int64_t n = 1;
    for (int64_t d : v.shape) n *= d;
    return n;
}

template <class T>
int read_classic(const NcFile* f, const Var& v, T* out) {
    const int ts = type_size(v.type);
    int64_t slab = 1;
    for (size_t k = v.record ? 1 : 0; k < v.shape.size(); ++k) slab *= v.shape[k];
    const int64_t nrec = v.record ? v.shape[0] : 1;
    std::vector<uint8_t> buf;
    const int64_t chunk = std::max<int64_t>(1, (int64_t(1) << 24) / ts);  // 16 MiB at a time
    for (int64_t r = 0; r < nrec; ++r) {
        const int64_t base = v.begin + (v.record ? r * f->recsize : 0);
        for (int64_t s = 0; s < slab; s += chunk) {
            const int64_t m = std::min(chunk, slab - s);
            buf.resize(size_t(m * ts));
            if (!pread_all(f->fd, buf.data(), m * ts, base + s * ts))
                return fail(-2, "read error in variable " + v.name + " of " + f->path);
            T* o = out + r * slab + s;
            for (int64_t i = 0; i < m; ++i) o[i] = be_value<T>(buf.data() + i * ts, v.type);
        }
    }
    return 0;
}

template <class T>
int read_var(const NcFile* f, const char* name, T* out, int64_t n, bool integer) {
    const Var* v = lookup(f, name);
    if (!v) return -3;
    if (n != count_of(*v))
        return fail(-3, "variable " + v->name + " has " + std::to_string(count_of(*v)) + " values, not " +
                            std::to_string(n));
    if (integer && !is_integer(v->type)) return fail(-3, "variable " + v->name + " is not an integer variable");
    if (v->type == NC_CHAR) return fail(-3, "variable " + v->name + " is a character variable");
    if (n == 0) return 0;
    if (f->format != 4) return read_classic(f, *v, out);
    Hdf5* H = hdf5();
    hid_t d = H->Dopen2(f->h5, name, kH5P_DEFAULT);
    if (d < 0) return fail(-2, "HDF5 could not open " + v->name);
    herr_t rc = H->Dread(d, integer ? H->native_llong : H->native_double, kH5S_ALL, kH5S_ALL, kH5P_DEFAULT, out);
    H->Dclose(d);
    return rc < 0 ? fail(-2, "HDF5 could not read " + v->name) : 0;
}

}  // namespace
}  // namespace dmx

using dmx::NcFile;

extern "C" {

int dmx_nc_open(const char* path, NcFile** out) {
    if (!path || !out) return dmx::fail(-1, "null argument");
    return dmx::open_file(path, out);
}

void dmx_nc_close(NcFile* f) { delete f; }

int dmx_nc_format(const NcFile* f) { return f ? f->format : -1; }

int64_t dmx_nc_num_vars(const NcFile* f) { return f ? int64_t(f->vars.size()) : -1; }

int dmx_nc_var_name(const NcFile* f, int64_t i, char* buf, int64_t cap) {
    if (!f || i < 0 || i >= int64_t(f->vars.size())) return dmx::fail(-1, "variable index out of range");
    const std::string& s = f->vars[size_t(i)].name;
    if (int64_t(s.size()) >= cap) return dmx::fail(-1, "name buffer too small");
    std::memcpy(buf, s.c_str(), s.size() + 1);
    return 0;
}

int dmx_nc_has_var(const NcFile* f, const char* name) { return f && f->find(name) >= 0 ? 1 : 0; }

int dmx_nc_var_info(const NcFile* f, const char* name, int* type, int* ndims, int64_t* shape, int max_dims) {
    const dmx::Var* v = dmx::lookup(f, name);
    if (!v) return -3;
    if (int(v->shape.size()) > max_dims) return dmx::fail(-1, "variable " + v->name + " has too many dimensions");
    *type = v->type;
    *ndims = int(v->shape.size());
    for (size_t k = 0; k < v->shape.size(); ++k) shape[k] = v->shape[k];
    return 0;
}

int dmx_nc_read_f64(const NcFile* f, const char* name, double* out, int64_t n) {
    return dmx::read_var(f, name, out, n, false);
}

int dmx_nc_read_i64(const NcFile* f, const char* name, int64_t* out, int64_t n) {
    return dmx::read_var(f, name, out, n, true);
}

int dmx_nc_global_text(const NcFile* f, const char* name, char* buf, int64_t cap, int64_t* len) {
    std::string s;
    bool found = false;
    if (f->format == 4) {
        s = dmx::read_string_attr(dmx::hdf5(), f->h5, name, &found);
    } else {
        for (const auto& a : f->gatts)
            if (a.name == name && a.type == dmx::NC_CHAR) {
                found = true;
                s.assign(a.raw.begin(), a.raw.end());
                s = s.substr(0, strnlen(s.c_str(), s.size()));
            }
    }
    if (!found) return 1;
    *len = int64_t(s.size());
    if (int64_t(s.size()) >= cap) return dmx::fail(-1, "attribute buffer too small");
    std::memcpy(buf, s.c_str(), s.size() + 1);
    return 0;
}

// sort_ring / dual_polygons (src/Grids/FileGrids.jl:22-41): every ring r, the CSR slice
// ids[offsets[r], offsets[r+1]) of point indices, is sorted in place counter-clockwise
// around a = normalize(axes[r]) seen from outside: by the angle
// atan(dot(cross(e1, v), a), dot(e1, v)) with e1 = points[ids[first]] - a and
// v = points[id] - a, stably, as Julia's sortperm.
int dmx_sort_rings(int64_t nrings, const int64_t* offsets, int64_t* ids, int64_t npoints, const double* points,
                   const double* axes) {
    int bad = 0;
#pragma omp parallel for schedule(static) reduction(| : bad)
    for (int64_t r = 0; r < nrings; ++r) {
        const int64_t b = offsets[r], e = offsets[r + 1];
        if (e <= b) continue;
        std::vector<std::pair<double, int64_t>> key;
        for (int64_t k = b; k < e; ++k)
            if (ids[k] < 0 || ids[k] >= npoints) bad = 1;
        if (bad) continue;
        const dmx::V3 a = dmx::normalize(dmx::load3(axes, r));
        const dmx::V3 e1 = dmx::load3(points, ids[b]) - a;
        for (int64_t k = b; k < e; ++k) {
            const dmx::V3 v = dmx::load3(points, ids[k]) - a;
            key.emplace_back(std::atan2(dmx::dot(dmx::cross(e1, v), a), dmx::dot(e1, v)), ids[k]);
        }
        std::stable_sort(key.begin(), key.end(),
                         [](const auto& x, const auto& y) { return x.first < y.first; });
        for (int64_t k = b; k < e; ++k) ids[k] = key[size_t(k - b)].second;
    }
    return bad ? dmx::fail(-3, "ring index out of range") : 0;
}

}  // extern "C"
