// HDF5 output through libhdf5 loaded at run time.  Only the C API entry
// points below are used; their prototypes are those of HDF5 1.10 (hid_t is
// a 64-bit integer, hsize_t an unsigned 64-bit integer).
#include "g4mi_h5.h"

#include <dlfcn.h>

#include <algorithm>
#include <cstdlib>
#include <mutex>

namespace g4mi {

namespace {

typedef int64_t hid_t;
typedef int herr_t;
typedef unsigned long long hsize_t;

constexpr unsigned kAccTrunc = 0x0002u, kAccRdonly = 0x0000u;
constexpr hid_t kDefault = 0;  // H5P_DEFAULT, H5S_ALL, H5E_DEFAULT
constexpr int kScalar = 0;     // H5S_SCALAR
constexpr int kNullTerm = 0;   // H5T_STR_NULLTERM

struct Api {
    herr_t (*open)();
    hid_t (*fcreate)(const char *, unsigned, hid_t, hid_t);
    hid_t (*fopen)(const char *, unsigned, hid_t);
    herr_t (*fclose)(hid_t);
    hid_t (*gcreate2)(hid_t, const char *, hid_t, hid_t, hid_t);
    hid_t (*gopen2)(hid_t, const char *, hid_t);
    herr_t (*gclose)(hid_t);
    int (*lexists)(hid_t, const char *, hid_t);
    hid_t (*screate)(int);
    hid_t (*screate_simple)(int, const hsize_t *, const hsize_t *);
    herr_t (*sclose)(hid_t);
    int (*sget_ndims)(hid_t);
    int (*sget_dims)(hid_t, hsize_t *, hsize_t *);
    hid_t (*dcreate2)(hid_t, const char *, hid_t, hid_t, hid_t, hid_t, hid_t);
    hid_t (*dopen2)(hid_t, const char *, hid_t);
    herr_t (*dwrite)(hid_t, hid_t, hid_t, hid_t, hid_t, const void *);
    herr_t (*dread)(hid_t, hid_t, hid_t, hid_t, hid_t, void *);
    hid_t (*dget_space)(hid_t);
    herr_t (*dclose)(hid_t);
    hid_t (*tcopy)(hid_t);
    herr_t (*tset_size)(hid_t, size_t);
    herr_t (*tset_strpad)(hid_t, int);
    herr_t (*tclose)(hid_t);
    hid_t (*acreate2)(hid_t, const char *, hid_t, hid_t, hid_t, hid_t);
    herr_t (*awrite)(hid_t, hid_t, const void *);
    herr_t (*aclose)(hid_t);
    herr_t (*eset_auto2)(hid_t, void *, void *);
    hid_t *native_double, *native_int, *c_s1;
};

std::once_flag g_once;
Api g_api;
bool g_ok = false;
std::string g_err;

void load_api() {
    const char *env = std::getenv("G4MI_HDF5");
    const char *names[] = {env, "libhdf5.so.103", "libhdf5.so", "/opt/conda/lib/libhdf5.so.103",
                           "/opt/conda/lib/libhdf5.so"};
    void *h = nullptr;
    for (const char *n : names)
        if (n && (h = dlopen(n, RTLD_NOW | RTLD_LOCAL))) break;
    if (!h) {
        g_err = "libhdf5 not found (set G4MI_HDF5 to its path)";
        return;
    }
    bool ok = true;
    auto sym = [&](const char *name) {
        void *p = dlsym(h, name);
        if (!p) {
            ok = false;
            g_err = std::string("libhdf5: missing symbol ") + name;
        }
        return p;
    };
    Api &a = g_api;
    a.open = (decltype(a.open))sym("H5open");
    a.fcreate = (decltype(a.fcreate))sym("H5Fcreate");
    a.fopen = (decltype(a.fopen))sym("H5Fopen");
    a.fclose = (decltype(a.fclose))sym("H5Fclose");
    a.gcreate2 = (decltype(a.gcreate2))sym("H5Gcreate2");
    a.gopen2 = (decltype(a.gopen2))sym("H5Gopen2");
    a.gclose = (decltype(a.gclose))sym("H5Gclose");
    a.lexists = (decltype(a.lexists))sym("H5Lexists");
    a.screate = (decltype(a.screate))sym("H5Screate");
    a.screate_simple = (decltype(a.screate_simple))sym("H5Screate_simple");
    a.sclose = (decltype(a.sclose))sym("H5Sclose");
    a.sget_ndims = (decltype(a.sget_ndims))sym("H5Sget_simple_extent_ndims");
    a.sget_dims = (decltype(a.sget_dims))sym("H5Sget_simple_extent_dims");
    a.dcreate2 = (decltype(a.dcreate2))sym("H5Dcreate2");
    a.dopen2 = (decltype(a.dopen2))sym("H5Dopen2");
    a.dwrite = (decltype(a.dwrite))sym("H5Dwrite");
    a.dread = (decltype(a.dread))sym("H5Dread");
    a.dget_space = (decltype(a.dget_space))sym("H5Dget_space");
    a.dclose = (decltype(a.dclose))sym("H5Dclose");
    a.tcopy = (decltype(a.tcopy))sym("H5Tcopy");
    a.tset_size = (decltype(a.tset_size))sym("H5Tset_size");
    a.tset_strpad = (decltype(a.tset_strpad))sym("H5Tset_strpad");
    a.tclose = (decltype(a.tclose))sym("H5Tclose");
    a.acreate2 = (decltype(a.acreate2))sym("H5Acreate2");
    a.awrite = (decltype(a.awrite))sym("H5Awrite");
    a.aclose = (decltype(a.aclose))sym("H5Aclose");
    a.eset_auto2 = (decltype(a.eset_auto2))sym("H5Eset_auto2");
    a.native_double = (hid_t *)sym("H5T_NATIVE_DOUBLE_g");
    a.native_int = (hid_t *)sym("H5T_NATIVE_INT_g");
    a.c_s1 = (hid_t *)sym("H5T_C_S1_g");
    if (!ok) return;
    if (a.open() < 0) {
        g_err = "H5open failed";
        return;
    }
    a.eset_auto2(kDefault, nullptr, nullptr);  // errors come back as return codes only
    g_ok = true;
}

const Api *api(std::string *err) {
    std::call_once(g_once, load_api);
    if (!g_ok) {
        if (err) *err = g_err;
        return nullptr;
    }
    return &g_api;
}

// "unit" attribute: fixed-length, null-terminated string
void write_unit(const Api *a, hid_t did, const std::string &unit) {
    const hid_t sid = a->screate(kScalar);
    const hid_t t = a->tcopy(*a->c_s1);
    a->tset_size(t, std::max<size_t>(unit.size(), 1));
    a->tset_strpad(t, kNullTerm);
    const hid_t att = a->acreate2(did, "unit", t, sid, kDefault, kDefault);
    if (att >= 0) {
        a->awrite(att, t, unit.c_str());
        a->aclose(att);
    }
    a->tclose(t);
    a->sclose(sid);
}

}  // namespace

bool h5_available(std::string *err) { return api(err) != nullptr; }

H5File::H5File(const std::string &path) {
    const Api *a = api(&err_);
    if (!a) return;
    fid_ = a->fcreate(path.c_str(), kAccTrunc, kDefault, kDefault);
    if (fid_ < 0) err_ = "cannot create " + path;
}

H5File::~H5File() { close(); }

void H5File::close() {
    const Api *a = api(nullptr);
    if (!a || fid_ < 0) return;
    for (auto it = groups_.rbegin(); it != groups_.rend(); ++it) a->gclose(it->second);
    groups_.clear();
    a->fclose(fid_);
    fid_ = -1;
}

int64_t H5File::group_of(const std::string &path, std::string *leaf) {
    const Api *a = api(nullptr);
    const size_t p = path.find_last_of('/');
    if (p == std::string::npos) {
        *leaf = path;
        return fid_;
    }
    *leaf = path.substr(p + 1);
    const std::string gpath = path.substr(0, p);
    for (auto &g : groups_)
        if (g.first == gpath) return g.second;
    // create or open each level
    hid_t cur = fid_;
    std::string sofar;
    size_t start = 0;
    while (start <= gpath.size()) {
        const size_t q = gpath.find('/', start);
        const std::string name = gpath.substr(start, q == std::string::npos ? std::string::npos : q - start);
        sofar += (sofar.empty() ? "" : "/") + name;
        hid_t g = -1;
        for (auto &e : groups_)
            if (e.first == sofar) g = e.second;
        if (g < 0) {
            g = a->lexists(cur, name.c_str(), kDefault) > 0 ? a->gopen2(cur, name.c_str(), kDefault)
                                                            : a->gcreate2(cur, name.c_str(), kDefault, kDefault, kDefault);
            if (g < 0) return -1;
            groups_.emplace_back(sofar, g);
        }
        cur = g;
        if (q == std::string::npos) break;
        start = q + 1;
    }
    return cur;
}

bool H5File::write_buffer(const std::string &path, const std::string &unit, const std::vector<double> &v,
                          size_t nrec, size_t ncol) {
    const Api *a = api(nullptr);
    if (!a || fid_ < 0 || v.size() != nrec * ncol || v.empty()) return false;
    std::string leaf;
    const hid_t g = group_of(path, &leaf);
    if (g < 0) return false;
    const hsize_t dims[2] = {(hsize_t)nrec, (hsize_t)ncol};
    const hid_t sid = a->screate_simple(2, dims, nullptr);
    const hid_t did = a->dcreate2(g, leaf.c_str(), *a->native_double, sid, kDefault, kDefault, kDefault);
    bool ok = did >= 0;
    if (ok) {
        write_unit(a, did, unit);
        ok = a->dwrite(did, *a->native_double, kDefault, kDefault, kDefault, v.data()) >= 0;
        a->dclose(did);
    }
    a->sclose(sid);
    return ok;
}

bool H5File::write_single(const std::string &path, const std::string &unit, const std::vector<double> &v) {
    const Api *a = api(nullptr);
    if (!a || fid_ < 0 || v.empty()) return false;
    std::string leaf;
    const hid_t g = group_of(path, &leaf);
    if (g < 0) return false;
    const hsize_t dims[1] = {(hsize_t)v.size()};
    const hid_t sid = a->screate_simple(1, dims, nullptr);
    const hid_t did = a->dcreate2(g, leaf.c_str(), *a->native_double, sid, kDefault, kDefault, kDefault);
    bool ok = did >= 0;
    if (ok) {
        write_unit(a, did, unit);
        ok = a->dwrite(did, *a->native_double, kDefault, kDefault, kDefault, v.data()) >= 0;
        a->dclose(did);
    }
    a->sclose(sid);
    return ok;
}

bool H5File::write_int(const std::string &path, const std::vector<int> &v) {
    const Api *a = api(nullptr);
    if (!a || fid_ < 0 || v.empty()) return false;
    std::string leaf;
    const hid_t g = group_of(path, &leaf);
    if (g < 0) return false;
    const hsize_t dims[1] = {(hsize_t)v.size()};
    const hid_t sid = a->screate_simple(1, dims, nullptr);
    const hid_t did = a->dcreate2(g, leaf.c_str(), *a->native_int, sid, kDefault, kDefault, kDefault);
    bool ok = did >= 0;
    if (ok) {
        ok = a->dwrite(did, *a->native_int, kDefault, kDefault, kDefault, v.data()) >= 0;
        a->dclose(did);
    }
    a->sclose(sid);
    return ok;
}

bool H5File::write_string(const std::string &path, const std::string &s) {
    const Api *a = api(nullptr);
    if (!a || fid_ < 0) return false;
    std::string leaf;
    const hid_t g = group_of(path, &leaf);
    if (g < 0) return false;
    const hsize_t dims[1] = {1};
    const hid_t sid = a->screate_simple(1, dims, nullptr);
    const hid_t t = a->tcopy(*a->c_s1);
    a->tset_size(t, std::max<size_t>(s.size(), 1));
    const hid_t did = a->dcreate2(g, leaf.c_str(), t, sid, kDefault, kDefault, kDefault);
    bool ok = did >= 0;
    if (ok) {
        std::string buf = s.empty() ? std::string(1, '\0') : s;
        ok = a->dwrite(did, t, kDefault, kDefault, kDefault, buf.data()) >= 0;
        a->dclose(did);
    }
    a->tclose(t);
    a->sclose(sid);
    return ok;
}

long long h5_read_doubles(const std::string &file, const std::string &dataset, std::vector<double> &out,
                          long long dims[2], std::string &err) {
    const Api *a = api(&err);
    if (!a) return -1;
    const hid_t fid = a->fopen(file.c_str(), kAccRdonly, kDefault);
    if (fid < 0) {
        err = "cannot open " + file;
        return -1;
    }
    long long n = -1;
    const hid_t did = a->dopen2(fid, dataset.c_str(), kDefault);
    if (did < 0) {
        err = "no dataset " + dataset + " in " + file;
    } else {
        const hid_t sid = a->dget_space(did);
        const int nd = a->sget_ndims(sid);
        hsize_t d[2] = {1, 1};
        if (nd < 1 || nd > 2) {
            err = "dataset " + dataset + " has rank " + std::to_string(nd);
        } else {
            a->sget_dims(sid, d, nullptr);
            if (nd == 1) d[1] = 1;
            out.assign((size_t)(d[0] * d[1]), 0.0);
            if (a->dread(did, *a->native_double, kDefault, kDefault, kDefault, out.data()) < 0) {
                err = "cannot read " + dataset;
            } else {
                dims[0] = (long long)d[0];
                dims[1] = nd == 1 ? 0 : (long long)d[1];
                n = (long long)(d[0] * d[1]);
            }
        }
        a->sclose(sid);
        a->dclose(did);
    }
    a->fclose(fid);
    return n;
}

}  // namespace g4mi
