// Restart and statistics files (src/LCSIO/save.jl, load.jl, hdf5.jl) through
// the HDF5 C library.  Host code only: fields and particles reach the host
// through liblcs_mi355x.so's public transfer calls.
#include "lcs_io.h"

#include <hdf5.h>
#include <unistd.h>

#include <cstring>
#include <string>
#include <vector>

namespace {

thread_local std::string g_err;

int fail(int code, const std::string &msg)
{
    g_err = msg;
    return code;
}

// The innermost message of the HDF5 error stack, for error texts.
herr_t walk_cb(unsigned n, const H5E_error2_t *e, void *data)
{
    auto *s = static_cast<std::string *>(data);
    if (n == 0 && e->desc) *s = e->desc;
    return 0;
}

std::string h5_reason()
{
    std::string s;
    H5Ewalk2(H5E_DEFAULT, H5E_WALK_UPWARD, walk_cb, &s);
    H5Eclear2(H5E_DEFAULT);
    return s.empty() ? std::string() : " (" + s + ")";
}

int h5_fail(const std::string &what) { return fail(LCS_ERR_STATE, what + h5_reason()); }

int core_fail(int rc, const char *call)
{
    return fail(rc, std::string(call) + ": " + lcs_last_error());
}

// hid_t owner closing with the matching H5?close
struct Hid {
    hid_t id = -1;
    herr_t (*close)(hid_t) = nullptr;
    Hid() = default;
    Hid(hid_t i, herr_t (*c)(hid_t)) : id(i), close(c) {}
    Hid(const Hid &) = delete;
    Hid &operator=(const Hid &) = delete;
    Hid(Hid &&o) noexcept : id(o.id), close(o.close) { o.id = -1; }
    Hid &operator=(Hid &&o) noexcept
    {
        if (this != &o) {
            if (id >= 0 && close) close(id);
            id = o.id;
            close = o.close;
            o.id = -1;
        }
        return *this;
    }
    ~Hid()
    {
        if (id >= 0 && close) close(id);
    }
    bool ok() const { return id >= 0; }
    operator hid_t() const { return id; }
};

// Errors are reported through lcs_io_last_error, not printed: the default
// error stack's printing is off for the duration of a call and restored after.
struct Quiet {
    H5E_auto2_t func = nullptr;
    void *data = nullptr;
    Quiet()
    {
        H5Eget_auto2(H5E_DEFAULT, &func, &data);
        H5Eset_auto2(H5E_DEFAULT, nullptr, nullptr);
    }
    ~Quiet() { H5Eset_auto2(H5E_DEFAULT, func, data); }
    Quiet(const Quiet &) = delete;
    Quiet &operator=(const Quiet &) = delete;
};

bool file_exists(const char *file) { return access(file, F_OK) == 0; }

// h5open(file, "cw"): open read-write, creating the file when absent.
int open_cw(const char *file, Hid *f)
{
    if (file_exists(file)) {
        if (H5Fis_hdf5(file) <= 0) return fail(LCS_ERR_INVALID, std::string(file) + " is not an HDF5 file");
        *f = Hid(H5Fopen(file, H5F_ACC_RDWR, H5P_DEFAULT), H5Fclose);
    } else {
        *f = Hid(H5Fcreate(file, H5F_ACC_EXCL, H5P_DEFAULT, H5P_DEFAULT), H5Fclose);
    }
    if (!f->ok()) return h5_fail(std::string("cannot open ") + file + " for writing");
    return LCS_OK;
}

int open_r(const char *file, Hid *f)
{
    if (!file_exists(file)) return fail(LCS_ERR_INVALID, std::string(file) + " does not exist");
    *f = Hid(H5Fopen(file, H5F_ACC_RDONLY, H5P_DEFAULT), H5Fclose);
    if (!f->ok()) return h5_fail(std::string("cannot open ") + file);
    return LCS_OK;
}

// true when every component of path names an existing link below loc
bool path_exists(hid_t loc, const std::string &path)
{
    if (path.empty() || path == "/") return true;
    size_t pos = 0;
    while (true) {
        size_t next = path.find('/', pos);
        std::string prefix = path.substr(0, next);
        if (!prefix.empty() && prefix != "/") {
            htri_t e = H5Lexists(loc, prefix.c_str(), H5P_DEFAULT);
            if (e <= 0) {
                H5Eclear2(H5E_DEFAULT);
                return false;
            }
        }
        if (next == std::string::npos) return true;
        pos = next + 1;
    }
}

// link creation as HDF5.jl does it: UTF-8 names, missing groups created
Hid link_create_plist()
{
    Hid lcpl(H5Pcreate(H5P_LINK_CREATE), H5Pclose);
    if (lcpl.ok()) {
        H5Pset_create_intermediate_group(lcpl, 1);
        H5Pset_char_encoding(lcpl, H5T_CSET_UTF8);
    }
    return lcpl;
}

hid_t file_float_type(int dtype) { return dtype == LCS_F32 ? H5T_IEEE_F32LE : H5T_IEEE_F64LE; }
hid_t mem_float_type(int dtype) { return dtype == LCS_F32 ? H5T_NATIVE_FLOAT : H5T_NATIVE_DOUBLE; }

const char *kFieldNames[4] = {"U", "V", "W", "P"};  // TAG_FLOW_* (src/LCSIO/LCSIO.jl)
const int kFieldIds[4] = {LCS_FIELD_U, LCS_FIELD_V, LCS_FIELD_W, LCS_FIELD_P};

// particle datasets in the order saverestart writes them (save.jl:47-53)
struct PropTag {
    const char *name;
    int prop;
};
const PropTag kPropTags[8] = {{"id", LCS_PROP_ID}, {"x", LCS_PROP_X}, {"y", LCS_PROP_Y}, {"z", LCS_PROP_Z},
                              {"u", LCS_PROP_U},   {"v", LCS_PROP_V}, {"w", LCS_PROP_W}, {"diam", LCS_PROP_DIAM}};

// memory dataspace of a whole field with halo, the interior selected
Hid interior_space(const lcs_box_desc &d)
{
    const hsize_t h = (hsize_t)d.halo;
    hsize_t full[3] = {(hsize_t)d.nz + 2 * h, (hsize_t)d.ny + 2 * h, (hsize_t)d.nx + 2 * h};
    Hid s(H5Screate_simple(3, full, nullptr), H5Sclose);
    if (!s.ok()) return s;
    hsize_t start[3] = {h, h, h};
    hsize_t count[3] = {(hsize_t)d.nz, (hsize_t)d.ny, (hsize_t)d.nx};
    if (H5Sselect_hyperslab(s, H5S_SELECT_SET, start, nullptr, count, nullptr) < 0) return Hid();
    return s;
}

size_t field_elems(const lcs_box_desc &d)
{
    return (size_t)(d.nx + 2 * d.halo) * (size_t)(d.ny + 2 * d.halo) * (size_t)(d.nz + 2 * d.halo);
}

int check_desc(lcs_box *box, lcs_box_desc *d, int *nprof)
{
    int rc = lcs_box_describe(box, d, nprof);
    if (rc) return core_fail(rc, "lcs_box_describe");
    if (d->dtype != LCS_F32 && d->dtype != LCS_F64) return fail(LCS_ERR_INVALID, "unknown box float type");
    return LCS_OK;
}

std::string dims_text(const int64_t *d, int n)
{
    std::string s = "(";
    for (int q = 0; q < n; q++) s += (q ? ", " : "") + std::to_string(d[q]);
    return s + (n == 1 ? ",)" : ")");
}

// write one dataset of n elements (rank 1) from host memory
int write_vector(hid_t g, const char *name, hid_t ftype, hid_t mtype, const void *data, hsize_t n,
                 const std::string &where)
{
    Hid sp(H5Screate_simple(1, &n, nullptr), H5Sclose);
    if (!sp.ok()) return h5_fail("dataspace for " + where);
    Hid lcpl = link_create_plist();
    Hid ds(H5Dcreate2(g, name, ftype, sp, lcpl, H5P_DEFAULT, H5P_DEFAULT), H5Dclose);
    if (!ds.ok()) return h5_fail("cannot create dataset " + where);
    if (H5Dwrite(ds, mtype, H5S_ALL, H5S_ALL, H5P_DEFAULT, data) < 0) return h5_fail("cannot write " + where);
    return LCS_OK;
}

}  // namespace

extern "C" {

const char *lcs_io_last_error(void) { return g_err.c_str(); }

int lcs_save_restart(lcs_box *box, const char *file)
{
    if (!box || !file) return fail(LCS_ERR_INVALID, "null argument");
    Quiet quiet;
    lcs_box_desc d;
    int nprof = 0, rc;
    if ((rc = check_desc(box, &d, &nprof))) return rc;
    Hid f;
    if ((rc = open_cw(file, &f))) return rc;

    // save_flow_fields (save.jl:58-92): one host copy of the whole field at a
    // time, the interior selected by the memory dataspace
    if (path_exists(f, "flow")) return fail(LCS_ERR_STATE, std::string(file) + " already holds group flow");
    Hid lcpl = link_create_plist();
    Hid g(H5Gcreate2(f, "flow", lcpl, H5P_DEFAULT, H5P_DEFAULT), H5Gclose);
    if (!g.ok()) return h5_fail("cannot create group flow");
    const size_t es = d.dtype == LCS_F32 ? 4 : 8;
    {
        std::vector<unsigned char> host(field_elems(d) * es);
        Hid mem = interior_space(d);
        hsize_t fdims[3] = {(hsize_t)d.nz, (hsize_t)d.ny, (hsize_t)d.nx};
        Hid fsp(H5Screate_simple(3, fdims, nullptr), H5Sclose);
        if (!mem.ok() || !fsp.ok()) return h5_fail("field dataspaces");
        for (int q = 0; q < 4; q++) {
            if ((rc = lcs_field_download(box, kFieldIds[q], host.data(), host.size())))
                return core_fail(rc, "lcs_field_download");
            std::string where = std::string("flow/") + kFieldNames[q];
            Hid ds(H5Dcreate2(g, kFieldNames[q], file_float_type(d.dtype), fsp, lcpl, H5P_DEFAULT, H5P_DEFAULT),
                   H5Dclose);
            if (!ds.ok()) return h5_fail("cannot create dataset " + where);
            if (H5Dwrite(ds, mem_float_type(d.dtype), mem, H5S_ALL, H5P_DEFAULT, host.data()) < 0)
                return h5_fail("cannot write " + where);
        }
    }

    // save_particle_{counts,id,position,velocity,size} (save.jl:94-167)
    for (int i = 0; i < nprof; i++) {
        lcs_particle_desc pd;
        if ((rc = lcs_particles_describe(box, i, &pd))) return core_fail(rc, "lcs_particles_describe");
        std::string gname = "particles/" + std::to_string(i + 1);
        Hid pg;
        if (path_exists(f, gname))
            pg = Hid(H5Gopen2(f, gname.c_str(), H5P_DEFAULT), H5Gclose);
        else
            pg = Hid(H5Gcreate2(f, gname.c_str(), lcpl, H5P_DEFAULT, H5P_DEFAULT), H5Gclose);
        if (!pg.ok()) return h5_fail("cannot create group " + gname);
        int64_t counts[1] = {pd.nvalid};  // one process: counts[rank + 1] = nvalid
        if ((rc = write_vector(pg, "counts", H5T_STD_I64LE, H5T_NATIVE_INT64, counts, 1, gname + "/counts")))
            return rc;
        if (pd.nvalid == 0) continue;  // write_property skips empty properties (hdf5.jl:36)
        std::vector<unsigned char> host((size_t)pd.ntotal * 8);
        for (const PropTag &t : kPropTags) {
            const bool id = t.prop == LCS_PROP_ID;
            const size_t pes = id ? 8 : es;
            if ((rc = lcs_particles_download(box, i, t.prop, host.data(), (size_t)pd.ntotal * pes)))
                return core_fail(rc, "lcs_particles_download");
            if ((rc = write_vector(pg, t.name, id ? H5T_STD_I64LE : file_float_type(d.dtype),
                                   id ? H5T_NATIVE_INT64 : mem_float_type(d.dtype), host.data(), (hsize_t)pd.nvalid,
                                   gname + "/" + t.name)))
                return rc;
        }
    }
    if (H5Fflush(f, H5F_SCOPE_LOCAL) < 0) return h5_fail(std::string("cannot flush ") + file);
    return LCS_OK;
}

int lcs_load_flow_fields(lcs_box *box, const char *file)
{
    if (!box || !file) return fail(LCS_ERR_INVALID, "null argument");
    Quiet quiet;
    lcs_box_desc d;
    int rc;
    if ((rc = check_desc(box, &d, nullptr))) return rc;
    Hid f;
    if ((rc = open_r(file, &f))) return rc;
    const size_t es = d.dtype == LCS_F32 ? 4 : 8;
    std::vector<unsigned char> host(field_elems(d) * es, 0);
    Hid mem = interior_space(d);
    if (!mem.ok()) return h5_fail("field dataspace");
    for (int q = 0; q < 4; q++) {
        std::string where = std::string("flow/") + kFieldNames[q];
        if (!path_exists(f, where)) return fail(LCS_ERR_INVALID, std::string(file) + " has no dataset " + where);
        Hid ds(H5Dopen2(f, where.c_str(), H5P_DEFAULT), H5Dclose);
        if (!ds.ok()) return h5_fail("cannot open " + where);
        Hid sp(H5Dget_space(ds), H5Sclose);
        int rank = H5Sget_simple_extent_ndims(sp);
        hsize_t dims[8] = {};
        if (rank < 0 || rank > 8 || H5Sget_simple_extent_dims(sp, dims, nullptr) < 0)
            return h5_fail("cannot read the shape of " + where);
        int64_t got[8], want[3] = {d.nx, d.ny, d.nz};
        for (int r = 0; r < rank; r++) got[r] = (int64_t)dims[rank - 1 - r];
        if (rank != 3 || got[0] != want[0] || got[1] != want[1] || got[2] != want[2])
            return fail(LCS_ERR_INVALID, std::string("dataset ") + kFieldNames[q] + " size mismatch, expected " +
                                             dims_text(want, 3) + " got " + dims_text(got, rank));
        if (H5Dread(ds, mem_float_type(d.dtype), mem, H5S_ALL, H5P_DEFAULT, host.data()) < 0)
            return h5_fail("cannot read " + where);
        if ((rc = lcs_field_upload(box, kFieldIds[q], host.data(), host.size())))
            return core_fail(rc, "lcs_field_upload");
    }
    if (d.halo > 0) {
        unsigned mask = (1u << LCS_FIELD_U) | (1u << LCS_FIELD_V) | (1u << LCS_FIELD_W) | (1u << LCS_FIELD_P);
        if ((rc = lcs_fill_ghost(box, mask, d.halo, 1))) return core_fail(rc, "lcs_fill_ghost");
        if ((rc = lcs_sync(box))) return core_fail(rc, "lcs_sync");
    }
    return LCS_OK;
}

int lcs_load_particles(lcs_box *box, const char *file, int profile)
{
    if (!box || !file) return fail(LCS_ERR_INVALID, "null argument");
    Quiet quiet;
    lcs_box_desc d;
    int nprof = 0, rc;
    if ((rc = check_desc(box, &d, &nprof))) return rc;
    if (profile < 0 || profile >= nprof) return fail(LCS_ERR_STATE, "no such particle profile");
    lcs_particle_desc pd;
    if ((rc = lcs_particles_describe(box, profile, &pd))) return core_fail(rc, "lcs_particles_describe");
    Hid f;
    if ((rc = open_r(file, &f))) return rc;
    std::string gname = "particles/" + std::to_string(profile + 1);
    if (!path_exists(f, gname + "/counts"))
        return fail(LCS_ERR_INVALID, std::string(file) + " has no dataset " + gname + "/counts");

    // load_particle_counts! (load.jl:73-90)
    std::string cw = gname + "/counts";
    Hid cds(H5Dopen2(f, cw.c_str(), H5P_DEFAULT), H5Dclose);
    if (!cds.ok()) return h5_fail("cannot open " + cw);
    Hid csp(H5Dget_space(cds), H5Sclose);
    hssize_t ncounts = H5Sget_simple_extent_npoints(csp);
    if (ncounts != 1)
        return fail(LCS_ERR_INVALID, "process count mismatch, expected 1 got " + std::to_string((long long)ncounts));
    int64_t nvalid = 0;
    if (H5Dread(cds, H5T_NATIVE_INT64, H5S_ALL, H5S_ALL, H5P_DEFAULT, &nvalid) < 0)
        return h5_fail("cannot read " + cw);
    if (nvalid != pd.nvalid)
        return fail(LCS_ERR_INVALID, "particle count mismatch: profile " + std::to_string(profile) + " holds " +
                                         std::to_string((long long)pd.nvalid) + " valid particles, " + file +
                                         " holds " + std::to_string((long long)nvalid));
    if (nvalid == 0) return LCS_OK;  // read_property skips empty properties (hdf5.jl:52)

    // load_particle_{id,position,velocity,size}! (load.jl:44-66,88-118): the
    // first nvalid elements; the invalid tail keeps what the profile holds
    const size_t es = d.dtype == LCS_F32 ? 4 : 8;
    std::vector<unsigned char> host((size_t)pd.ntotal * 8);
    for (const PropTag &t : kPropTags) {
        const bool id = t.prop == LCS_PROP_ID;
        const size_t bytes = (size_t)pd.ntotal * (id ? 8 : es);
        std::string where = gname + "/" + t.name;
        if (!path_exists(f, where)) return fail(LCS_ERR_INVALID, std::string(file) + " has no dataset " + where);
        Hid ds(H5Dopen2(f, where.c_str(), H5P_DEFAULT), H5Dclose);
        if (!ds.ok()) return h5_fail("cannot open " + where);
        Hid sp(H5Dget_space(ds), H5Sclose);
        hssize_t npts = H5Sget_simple_extent_npoints(sp);
        if (npts != nvalid)
            return fail(LCS_ERR_INVALID, std::string("dataset ") + t.name + " size mismatch, expected " +
                                             std::to_string((long long)nvalid) + " got " +
                                             std::to_string((long long)npts));
        if ((rc = lcs_particles_download(box, profile, t.prop, host.data(), bytes)))
            return core_fail(rc, "lcs_particles_download");
        hsize_t n = (hsize_t)nvalid;
        Hid mem(H5Screate_simple(1, &n, nullptr), H5Sclose);
        if (H5Dread(ds, id ? H5T_NATIVE_INT64 : mem_float_type(d.dtype), mem, H5S_ALL, H5P_DEFAULT, host.data()) < 0)
            return h5_fail("cannot read " + where);
        if ((rc = lcs_particles_upload(box, profile, t.prop, host.data(), bytes)))
            return core_fail(rc, "lcs_particles_upload");
    }
    return LCS_OK;
}

}  // extern "C"

namespace {

int check_entries(const lcs_h5_entry *entries, int n)
{
    if (n < 0 || (n > 0 && !entries)) return fail(LCS_ERR_INVALID, "null argument");
    for (int q = 0; q < n; q++) {
        const lcs_h5_entry &e = entries[q];
        if (!e.path || !e.path[0] || e.path[0] == '/')
            return fail(LCS_ERR_INVALID, "entry " + std::to_string(q) + ": path must be relative and non-empty");
        std::string at = std::string("entry ") + e.path + ": ";
        if (e.type < LCS_H5_I64 || e.type > LCS_H5_STR) return fail(LCS_ERR_INVALID, at + "unknown type");
        if (e.rank < 0 || e.rank > 3) return fail(LCS_ERR_INVALID, at + "rank must be 0 to 3");
        if (e.type == LCS_H5_STR && e.rank != 0) return fail(LCS_ERR_INVALID, at + "strings are scalars");
        int64_t elems = 1;
        for (int r = 0; r < e.rank; r++) {
            if (e.dims[r] < 0) return fail(LCS_ERR_INVALID, at + "negative size");
            elems *= e.dims[r];
        }
        if (e.type == LCS_H5_STR && e.dims[0] < 0) return fail(LCS_ERR_INVALID, at + "negative string size");
        if (elems > 0 && !e.data && !(e.type == LCS_H5_STR && e.dims[0] == 0))
            return fail(LCS_ERR_INVALID, at + "no data");
    }
    return LCS_OK;
}

// writes every entry below loc; names in messages are prefixed with where
int write_entries(hid_t loc, const std::string &where, const lcs_h5_entry *entries, int n)
{
    Hid lcpl = link_create_plist();
    for (int q = 0; q < n; q++) {
        const lcs_h5_entry &e = entries[q];
        hsize_t hd[3];
        bool empty = false;
        for (int r = 0; r < e.rank; r++) {
            hd[e.rank - 1 - r] = (hsize_t)e.dims[r];  // Julia order -> HDF5 order
            empty |= e.dims[r] == 0;
        }
        if (empty) continue;
        std::string at = where + e.path;
        if (path_exists(loc, e.path)) return fail(LCS_ERR_STATE, at + " already exists");
        hid_t ftype, mtype;
        Hid strtype;
        switch (e.type) {
        case LCS_H5_I64: ftype = H5T_STD_I64LE; mtype = H5T_NATIVE_INT64; break;
        case LCS_H5_F64: ftype = H5T_IEEE_F64LE; mtype = H5T_NATIVE_DOUBLE; break;
        case LCS_H5_F32: ftype = H5T_IEEE_F32LE; mtype = H5T_NATIVE_FLOAT; break;
        default:
            strtype = Hid(H5Tcopy(H5T_C_S1), H5Tclose);
            if (!strtype.ok() || H5Tset_size(strtype, (size_t)(e.dims[0] > 0 ? e.dims[0] : 1)) < 0 ||
                H5Tset_cset(strtype, H5T_CSET_UTF8) < 0 || H5Tset_strpad(strtype, H5T_STR_NULLPAD) < 0)
                return h5_fail("string type for " + at);
            ftype = mtype = strtype;
        }
        Hid sp(e.rank == 0 ? H5Screate(H5S_SCALAR) : H5Screate_simple(e.rank, hd, nullptr), H5Sclose);
        if (!sp.ok()) return h5_fail("dataspace for " + at);
        Hid ds(H5Dcreate2(loc, e.path, ftype, sp, lcpl, H5P_DEFAULT, H5P_DEFAULT), H5Dclose);
        if (!ds.ok()) return h5_fail("cannot create dataset " + at);
        const char zero = 0;
        const void *src = (e.type == LCS_H5_STR && e.dims[0] == 0) ? &zero : e.data;
        if (H5Dwrite(ds, mtype, H5S_ALL, H5S_ALL, H5P_DEFAULT, src) < 0) return h5_fail("cannot write " + at);
    }
    return LCS_OK;
}

}  // namespace

extern "C" {

int lcs_save_stat(const char *file, int64_t step, const lcs_h5_entry *entries, int n)
{
    if (!file) return fail(LCS_ERR_INVALID, "null argument");
    int rc;
    if ((rc = check_entries(entries, n))) return rc;
    Quiet quiet;
    Hid f;
    if ((rc = open_cw(file, &f))) return rc;
    const std::string key = std::to_string((long long)step);
    // haskey(h, key) && delete_object(h, key) (save.jl:24-26)
    if (path_exists(f, key) && H5Ldelete(f, key.c_str(), H5P_DEFAULT) < 0) return h5_fail("cannot delete " + key);
    Hid lcpl = link_create_plist();
    Hid g(H5Gcreate2(f, key.c_str(), lcpl, H5P_DEFAULT, H5P_DEFAULT), H5Gclose);
    if (!g.ok()) return h5_fail("cannot create group " + key);
    if ((rc = write_entries(g, key + "/", entries, n))) return rc;
    if (H5Fflush(f, H5F_SCOPE_LOCAL) < 0) return h5_fail(std::string("cannot flush ") + file);
    return LCS_OK;
}

int lcs_h5_write(const char *file, const lcs_h5_entry *entries, int n)
{
    if (!file) return fail(LCS_ERR_INVALID, "null argument");
    int rc;
    if ((rc = check_entries(entries, n))) return rc;
    Quiet quiet;
    Hid f;
    if ((rc = open_cw(file, &f))) return rc;
    if ((rc = write_entries(f, "", entries, n))) return rc;
    if (H5Fflush(f, H5F_SCOPE_LOCAL) < 0) return h5_fail(std::string("cannot flush ") + file);
    return LCS_OK;
}

int lcs_h5_info(const char *file, const char *path, int *kind, int *type, int *rank, int64_t *dims)
{
    if (!file || !path || !kind) return fail(LCS_ERR_INVALID, "null argument");
    Quiet quiet;
    Hid f;
    int rc;
    if ((rc = open_r(file, &f))) return rc;
    *kind = LCS_H5_NONE;
    if (!path_exists(f, path)) return LCS_OK;
    H5O_info_t info;
    if (H5Oget_info_by_name(f, path, &info, H5P_DEFAULT) < 0) return h5_fail(std::string("cannot inspect ") + path);
    if (info.type == H5O_TYPE_GROUP) {
        *kind = LCS_H5_GROUP;
        return LCS_OK;
    }
    if (info.type != H5O_TYPE_DATASET) return LCS_OK;
    *kind = LCS_H5_DATASET;
    Hid ds(H5Dopen2(f, path, H5P_DEFAULT), H5Dclose);
    if (!ds.ok()) return h5_fail(std::string("cannot open ") + path);
    if (type) {
        Hid t(H5Dget_type(ds), H5Tclose);
        H5T_class_t c = H5Tget_class(t);
        size_t sz = H5Tget_size(t);
        *type = c == H5T_INTEGER ? LCS_H5_I64
                : c == H5T_FLOAT ? (sz == 4 ? LCS_H5_F32 : LCS_H5_F64)
                : c == H5T_STRING ? LCS_H5_STR
                                  : -1;
    }
    if (rank || dims) {
        Hid sp(H5Dget_space(ds), H5Sclose);
        int r = H5Sget_simple_extent_ndims(sp);
        hsize_t hd[8] = {};
        if (r < 0 || r > 8 || H5Sget_simple_extent_dims(sp, hd, nullptr) < 0)
            return h5_fail(std::string("cannot read the shape of ") + path);
        if (rank) *rank = r;
        if (dims)
            for (int q = 0; q < r; q++) dims[q] = (int64_t)hd[r - 1 - q];
    }
    return LCS_OK;
}

int lcs_h5_read(const char *file, const char *path, int type, void *out, int64_t capacity)
{
    if (!file || !path || !out) return fail(LCS_ERR_INVALID, "null argument");
    if (type < LCS_H5_I64 || type > LCS_H5_STR) return fail(LCS_ERR_INVALID, "unknown type");
    Quiet quiet;
    Hid f;
    int rc;
    if ((rc = open_r(file, &f))) return rc;
    if (!path_exists(f, path)) return fail(LCS_ERR_INVALID, std::string(file) + " has no dataset " + path);
    Hid ds(H5Dopen2(f, path, H5P_DEFAULT), H5Dclose);
    if (!ds.ok()) return h5_fail(std::string("cannot open dataset ") + path);
    Hid sp(H5Dget_space(ds), H5Sclose);
    hssize_t npts = H5Sget_simple_extent_npoints(sp);
    if (npts < 0) return h5_fail(std::string("cannot size ") + path);
    if (type != LCS_H5_STR) {
        if (npts > capacity)
            return fail(LCS_ERR_CAPACITY, std::string(path) + " holds " + std::to_string((long long)npts) +
                                              " elements, the buffer " + std::to_string((long long)capacity));
        hid_t mt = type == LCS_H5_I64 ? H5T_NATIVE_INT64 : type == LCS_H5_F64 ? H5T_NATIVE_DOUBLE : H5T_NATIVE_FLOAT;
        if (npts > 0 && H5Dread(ds, mt, H5S_ALL, H5S_ALL, H5P_DEFAULT, out) < 0)
            return h5_fail(std::string("cannot read ") + path);
        return LCS_OK;
    }
    Hid ft(H5Dget_type(ds), H5Tclose);
    if (H5Tget_class(ft) != H5T_STRING || npts != 1)
        return fail(LCS_ERR_INVALID, std::string(path) + " is not a scalar string");
    char *dst = static_cast<char *>(out);
    if (H5Tis_variable_str(ft) > 0) {
        Hid mt(H5Tcopy(H5T_C_S1), H5Tclose);
        H5Tset_size(mt, H5T_VARIABLE);
        H5Tset_cset(mt, H5Tget_cset(ft));
        char *s = nullptr;
        if (H5Dread(ds, mt, H5S_ALL, H5S_ALL, H5P_DEFAULT, &s) < 0) return h5_fail(std::string("cannot read ") + path);
        size_t len = s ? strlen(s) : 0;
        if ((int64_t)len + 1 > capacity) {
            H5free_memory(s);
            return fail(LCS_ERR_CAPACITY, std::string(path) + " needs " + std::to_string(len + 1) + " bytes");
        }
        if (len) memcpy(dst, s, len);
        dst[len] = 0;
        H5free_memory(s);
        return LCS_OK;
    }
    size_t sz = H5Tget_size(ft);
    std::vector<char> tmp(sz + 1, 0);
    if (H5Dread(ds, ft, H5S_ALL, H5S_ALL, H5P_DEFAULT, tmp.data()) < 0)
        return h5_fail(std::string("cannot read ") + path);
    size_t len = strnlen(tmp.data(), sz);
    if ((int64_t)len + 1 > capacity)
        return fail(LCS_ERR_CAPACITY, std::string(path) + " needs " + std::to_string(len + 1) + " bytes");
    memcpy(dst, tmp.data(), len);
    dst[len] = 0;
    return LCS_OK;
}

namespace {
herr_t key_cb(hid_t, const char *name, const H5L_info_t *, void *data)
{
    static_cast<std::vector<std::string> *>(data)->push_back(name);
    return 0;
}
}  // namespace

int lcs_h5_keys(const char *file, const char *path, char *names, size_t len, int *n, size_t *needed)
{
    if (!file || !path || !n || (len && !names)) return fail(LCS_ERR_INVALID, "null argument");
    Quiet quiet;
    Hid f;
    int rc;
    if ((rc = open_r(file, &f))) return rc;
    if (!path_exists(f, path)) return fail(LCS_ERR_INVALID, std::string(file) + " has no group " + path);
    Hid g(H5Gopen2(f, path, H5P_DEFAULT), H5Gclose);
    if (!g.ok()) return h5_fail(std::string("cannot open group ") + path);
    std::vector<std::string> keys;
    hsize_t idx = 0;
    if (H5Literate(g, H5_INDEX_NAME, H5_ITER_INC, &idx, key_cb, &keys) < 0)
        return h5_fail(std::string("cannot list ") + path);
    size_t total = 0;
    for (const auto &k : keys) total += k.size() + 1;
    *n = (int)keys.size();
    if (needed) *needed = total;
    if (total > len) return fail(LCS_ERR_CAPACITY, "key buffer needs " + std::to_string(total) + " bytes");
    size_t off = 0;
    for (const auto &k : keys) {
        memcpy(names + off, k.c_str(), k.size() + 1);
        off += k.size() + 1;
    }
    return LCS_OK;
}

}  // extern "C"
