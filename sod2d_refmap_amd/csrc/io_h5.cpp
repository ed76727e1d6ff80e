// libsod2d_io: SOD2D's partitioned HDF5 mesh files and restart files on the
// host, through the HDF5 C library (serial HDF5: one process reads or writes
// a partition's slices one after the other).
//
// Every dataset has the name, element type and shape that the reference's
// Fortran HDF5 layer gives it (src/lib_sod2d/sources/mod_hdf5.f90); the
// routine each function stands in for is cited at its declaration in
// include/sod2d_io.h.  Layout facts relied on here:
//   * the Fortran layer declares a dataset with Fortran dims (d1, d2), which
//     HDF5 stores as C dims (d2, d1): /Coords/Points, Fortran (3, N), is
//     C (N, 3), one row of x, y, z per node (mod_hdf5.f90:218-226, 1625-1657);
//   * connectivity datasets are 1-D, element-major: the nnode ids of an
//     element are contiguous (mod_hdf5.f90:887-910, 2139-2160);
//   * integers are 32-bit (H5T_NATIVE_INTEGER) except the 64-bit global node
//     ids; the uint8 flag isLinealOutput; coordinates are FP64;
//   * restart fields, and the restart time, are FP32 (rp_vtk = 4,
//     mod_constants.f90:6; mod_hdf5.f90:2887-2919, 3130-3212).
#include <hdf5.h>

#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "sod2d_io.h"

namespace {

thread_local std::string g_err;

int fail(int code, const char* fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    g_err = buf;
    return code;
}

// An HDF5 identifier closed when it goes out of scope.
class H5Id {
public:
    H5Id() = default;
    H5Id(hid_t id, herr_t (*close)(hid_t)) : id_(id), close_(close) {}
    H5Id(const H5Id&) = delete;
    H5Id& operator=(const H5Id&) = delete;
    ~H5Id() {
        if (id_ >= 0 && close_) close_(id_);
    }
    hid_t get() const { return id_; }
    bool ok() const { return id_ >= 0; }

private:
    hid_t id_ = -1;
    herr_t (*close_)(hid_t) = nullptr;
};

// HDF5 prints its error stack by default; failures are reported through the
// return codes instead.
void quiet() { H5Eset_auto2(H5E_DEFAULT, nullptr, nullptr); }

// Whether the link `path` exists; H5Lexists needs every parent to exist, so
// the components are checked in turn.
bool exists(hid_t f, const char* path) {
    const std::string p(path);
    size_t pos = p[0] == '/' ? 1 : 0;
    while (true) {
        const size_t nx = p.find('/', pos);
        const std::string cur = p.substr(0, nx);
        if (H5Lexists(f, cur.c_str(), H5P_DEFAULT) <= 0) return false;
        if (nx == std::string::npos) return true;
        pos = nx + 1;
    }
}

int extent(hid_t f, const char* name, int* rank, hsize_t dims[2]) {
    if (!exists(f, name)) return fail(SOD_IO_ERR_FORMAT, "dataset %s is missing", name);
    H5Id ds(H5Dopen2(f, name, H5P_DEFAULT), H5Dclose);
    if (!ds.ok()) return fail(SOD_IO_ERR_FORMAT, "cannot open dataset %s", name);
    H5Id sp(H5Dget_space(ds.get()), H5Sclose);
    const int r = H5Sget_simple_extent_ndims(sp.get());
    if (r < 1 || r > 2) return fail(SOD_IO_ERR_FORMAT, "dataset %s has rank %d", name, r);
    H5Sget_simple_extent_dims(sp.get(), dims, nullptr);
    *rank = r;
    return SOD_IO_OK;
}

int length1(hid_t f, const char* name, long long* n) {
    int r = 0;
    hsize_t d[2] = {0, 0};
    if (int e = extent(f, name, &r, d)) return e;
    if (r != 1) return fail(SOD_IO_ERR_FORMAT, "dataset %s is not 1-D", name);
    *n = (long long)d[0];
    return SOD_IO_OK;
}

// Rows [off, off+count) of a 1-D dataset (cols == 0) or of a 2-D dataset
// with `cols` columns, read into (or written from) `buf` as `memtype`.
int slab(hid_t f, const char* name, hid_t memtype, hsize_t off, hsize_t count, hsize_t cols, void* buf,
         bool write) {
    if (count == 0) return SOD_IO_OK;
    int r = 0;
    hsize_t d[2] = {0, 0};
    if (int e = extent(f, name, &r, d)) return e;
    if ((cols == 0 && r != 1) || (cols > 0 && (r != 2 || d[1] != cols)))
        return fail(SOD_IO_ERR_FORMAT, "dataset %s has an unexpected shape", name);
    if (off + count > d[0])
        return fail(SOD_IO_ERR_FORMAT, "dataset %s: rows %llu..%llu lie beyond its %llu rows", name,
                    (unsigned long long)off, (unsigned long long)(off + count), (unsigned long long)d[0]);
    H5Id ds(H5Dopen2(f, name, H5P_DEFAULT), H5Dclose);
    H5Id fs(H5Dget_space(ds.get()), H5Sclose);
    const hsize_t start[2] = {off, 0}, cnt[2] = {count, cols};
    if (H5Sselect_hyperslab(fs.get(), H5S_SELECT_SET, start, nullptr, cnt, nullptr) < 0)
        return fail(SOD_IO_ERR_FORMAT, "dataset %s: bad selection", name);
    H5Id ms(H5Screate_simple(r, cnt, nullptr), H5Sclose);
    const herr_t st = write ? H5Dwrite(ds.get(), memtype, ms.get(), fs.get(), H5P_DEFAULT, buf)
                            : H5Dread(ds.get(), memtype, ms.get(), fs.get(), H5P_DEFAULT, buf);
    if (st < 0) return fail(SOD_IO_ERR_FORMAT, "dataset %s: %s failed", name, write ? "write" : "read");
    return SOD_IO_OK;
}

int read_rows(hid_t f, const char* name, hid_t memtype, long long off, long long count, void* buf) {
    return slab(f, name, memtype, (hsize_t)off, (hsize_t)count, 0, buf, false);
}

int write_rows(hid_t f, const char* name, hid_t memtype, long long off, long long count, const void* buf) {
    return slab(f, name, memtype, (hsize_t)off, (hsize_t)count, 0, const_cast<void*>(buf), true);
}

int create_group(hid_t f, const char* name) {
    H5Id g(H5Gcreate2(f, name, H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT), H5Gclose);
    if (!g.ok()) return fail(SOD_IO_ERR_FILE, "cannot create group %s", name);
    return SOD_IO_OK;
}

int create_dataset(hid_t f, const char* name, hid_t type, int rank, const hsize_t* dims) {
    H5Id sp(H5Screate_simple(rank, dims, nullptr), H5Sclose);
    if (!sp.ok()) return fail(SOD_IO_ERR_FILE, "cannot make the dataspace of %s", name);
    H5Id ds(H5Dcreate2(f, name, type, sp.get(), H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT), H5Dclose);
    if (!ds.ok()) return fail(SOD_IO_ERR_FILE, "cannot create dataset %s", name);
    return SOD_IO_OK;
}

int create1(hid_t f, const char* name, hid_t type, long long n) {
    const hsize_t d = (hsize_t)n;
    return create_dataset(f, name, type, 1, &d);
}

// A per-partition vector of the file (one entry per partition) and the
// entry and prefix sum of partition `rank`.
int per_rank(hid_t f, const char* name, int nranks, int rank, long long* mine, long long* before) {
    std::vector<long long> v(nranks);
    if (int e = read_rows(f, name, H5T_NATIVE_LLONG, 0, nranks, v.data())) return e;
    long long s = 0;
    for (int r = 0; r < rank; ++r) s += v[r];
    *mine = v[rank];
    if (before) *before = s;
    if (v[rank] < 0) return fail(SOD_IO_ERR_FORMAT, "%s[%d] is negative", name, rank);
    return SOD_IO_OK;
}

// Element-major file rows -> Fortran (nelem, nnode) column-major, and back.
void elem_major_to_fortran(const std::vector<int>& buf, long long ne, int nnode, int* out) {
    for (long long e = 0; e < ne; ++e)
        for (int m = 0; m < nnode; ++m) out[e + ne * m] = buf[e * nnode + m];
}

void fortran_to_elem_major(const int* in, long long ne, int nnode, std::vector<int>& buf) {
    buf.resize((size_t)(ne * nnode));
    for (long long e = 0; e < ne; ++e)
        for (int m = 0; m < nnode; ++m) buf[e * nnode + m] = in[e + ne * m];
}

int append_name(std::string s, char* out, int cap) {
    if (out == nullptr || cap <= (int)s.size()) return fail(SOD_IO_ERR_ARG, "output buffer too small for %s", s.c_str());
    std::memcpy(out, s.c_str(), s.size() + 1);
    return (int)s.size();
}

}  // namespace

// ------------------------------------------------------------------ mesh

struct sod_h5_mesh {
    hid_t file = -1;
    long long info[SOD_H5_INFO_COUNT] = {};
    long long work_off = 0, per_off = 0, rcomm_off = 0, ncomm_off = 0;
    std::string coords;
};

struct sod_h5_mesh_writer {
    hid_t file = -1;
    int nranks = 0;
    bool periodic = false;
    std::vector<int> ne, nw, nper, ncr, ncn;
    std::vector<long long> nn;
    std::vector<char> done;
    long long pre(const std::vector<int>& v, int r) const {
        long long s = 0;
        for (int i = 0; i < r; ++i) s += v[i];
        return s;
    }
    long long pre(const std::vector<long long>& v, int r) const {
        long long s = 0;
        for (int i = 0; i < r; ++i) s += v[i];
        return s;
    }
};

extern "C" {

const char* sod_io_last_error(void) { return g_err.c_str(); }

int sod_h5_mesh_file_name(const char* path, const char* name, int nranks, char* out, int cap) {
    if (path == nullptr || name == nullptr || nranks < 1) return fail(SOD_IO_ERR_ARG, "mesh_file_name: bad argument");
    return append_name(std::string(path) + name + "-" + std::to_string(nranks) + ".hdf", out, cap);
}

int sod_h5_restart_file_name(const char* path, const char* mesh_name, int nranks, int restart_cnt, char* out,
                             int cap) {
    if (path == nullptr || mesh_name == nullptr || nranks < 1)
        return fail(SOD_IO_ERR_ARG, "restart_file_name: bad argument");
    return append_name(std::string(path) + "restart_" + mesh_name + "-" + std::to_string(nranks) + "_" +
                           std::to_string(restart_cnt) + ".h5",
                       out, cap);
}

int sod_h5_mesh_open(const char* path, int rank, sod_h5_mesh** out) {
    if (out == nullptr || path == nullptr) return fail(SOD_IO_ERR_ARG, "mesh_open: NULL argument");
    *out = nullptr;
    quiet();
    const hid_t f = H5Fopen(path, H5F_ACC_RDONLY, H5P_DEFAULT);
    if (f < 0) return fail(SOD_IO_ERR_FILE, "cannot open mesh file %s", path);
    sod_h5_mesh* m = new sod_h5_mesh();
    m->file = f;
    long long* I = m->info;
    auto bail = [&](int code) {
        sod_h5_mesh_close(m);
        return code;
    };
    if (int e = length1(f, "/globalIds/globalIdSrl", &I[SOD_H5_TOTAL_NODES])) return bail(e);
    if (int e = length1(f, "/globalIds/elemGid", &I[SOD_H5_TOTAL_ELEMS])) return bail(e);
    unsigned char lin = 0;
    if (int e = read_rows(f, "/meshOutputInfo/isLinealOutput", H5T_NATIVE_UCHAR, 0, 1, &lin)) return bail(e);
    I[SOD_H5_IS_LINEAL_OUTPUT] = lin == 1;
    int porder = 0;
    if (int e = read_rows(f, "/order/porder", H5T_NATIVE_INT, 0, 1, &porder)) return bail(e);
    I[SOD_H5_PORDER] = porder;
    if (porder != 3)
        return bail(fail(SOD_IO_ERR_PORDER, "mesh order %d in %s; this build is fixed at p = 3", porder, path));
    I[SOD_H5_NNODE] = 64;
    long long nranks = 0;
    if (int e = length1(f, "/Connectivity/numWorkingNodesRankPar", &nranks)) return bail(e);
    if (nranks < 1 || rank < 0 || rank >= nranks)
        return bail(fail(SOD_IO_ERR_ARG, "partition %d requested from a file of %lld partitions", rank, nranks));
    I[SOD_H5_NRANKS] = nranks;
    const int R = (int)nranks;
    if (R >= 2) {
        long long v = 0;
        if (int e = per_rank(f, "/Parallel_data/rankNodeStart", R, rank, &I[SOD_H5_RANK_NODE_START], nullptr)) return bail(e);
        if (int e = per_rank(f, "/Parallel_data/rankNodeEnd", R, rank, &I[SOD_H5_RANK_NODE_END], nullptr)) return bail(e);
        if (int e = per_rank(f, "/Parallel_data/rankElemStart", R, rank, &I[SOD_H5_RANK_ELEM_START], nullptr)) return bail(e);
        if (int e = per_rank(f, "/Parallel_data/rankElemEnd", R, rank, &I[SOD_H5_RANK_ELEM_END], nullptr)) return bail(e);
        if (int e = per_rank(f, "/Parallel_data/numRanksWithComms", R, rank, &v, &m->rcomm_off)) return bail(e);
        I[SOD_H5_NUM_RANKS_WITH_COMMS] = v;
        if (int e = per_rank(f, "/Parallel_data/numNodesToComm", R, rank, &v, &m->ncomm_off)) return bail(e);
        I[SOD_H5_NUM_NODES_TO_COMM] = v;
    } else {
        I[SOD_H5_RANK_NODE_START] = 1;
        I[SOD_H5_RANK_NODE_END] = I[SOD_H5_TOTAL_NODES];
        I[SOD_H5_RANK_ELEM_START] = 1;
        I[SOD_H5_RANK_ELEM_END] = I[SOD_H5_TOTAL_ELEMS];
    }
    I[SOD_H5_NUM_NODES_RANK] = I[SOD_H5_RANK_NODE_END] - I[SOD_H5_RANK_NODE_START] + 1;
    I[SOD_H5_NUM_ELEMS_RANK] = I[SOD_H5_RANK_ELEM_END] - I[SOD_H5_RANK_ELEM_START] + 1;
    if (I[SOD_H5_RANK_NODE_START] < 1 || I[SOD_H5_NUM_NODES_RANK] < 0 ||
        I[SOD_H5_RANK_NODE_END] > I[SOD_H5_TOTAL_NODES] || I[SOD_H5_RANK_ELEM_START] < 1 ||
        I[SOD_H5_NUM_ELEMS_RANK] < 0 || I[SOD_H5_RANK_ELEM_END] > I[SOD_H5_TOTAL_ELEMS])
        return bail(fail(SOD_IO_ERR_FORMAT, "partition %d: node or element range outside the file", rank));
    if (int e = per_rank(f, "/Connectivity/numWorkingNodesRankPar", R, rank, &I[SOD_H5_NUM_WORKING_NODES],
                         &m->work_off))
        return bail(e);
    I[SOD_H5_IS_PERIODIC] = exists(f, "/Periodic_data");
    if (I[SOD_H5_IS_PERIODIC]) {
        if (int e = per_rank(f, "/Periodic_data/nPerRankPar", R, rank, &I[SOD_H5_NPER], &m->per_off)) return bail(e);
    }
    I[SOD_H5_HAS_BOUNDARIES] = exists(f, "/Boundary_data");
    m->coords = I[SOD_H5_IS_LINEAL_OUTPUT] ? "/VTKHDF/Points" : "/Coords/Points";
    *out = m;
    return SOD_IO_OK;
}

int sod_h5_mesh_info(const sod_h5_mesh* m, long long* info) {
    if (m == nullptr || info == nullptr) return fail(SOD_IO_ERR_ARG, "mesh_info: NULL argument");
    std::memcpy(info, m->info, sizeof(m->info));
    return SOD_IO_OK;
}

int sod_h5_mesh_read(const sod_h5_mesh* m, int* connecParOrig, int* connecParWork, int* workingNodesPar,
                     double* coordPar, int* masSlaRankPar, long long* globalIdSrl, long long* globalIdPar,
                     int* elemGid, int* a2ijk, int* gmsh2ijk, int* vtk2ijk) {
    if (m == nullptr) return fail(SOD_IO_ERR_ARG, "mesh_read: NULL handle");
    const hid_t f = m->file;
    const long long* I = m->info;
    const long long ne = I[SOD_H5_NUM_ELEMS_RANK], nn = I[SOD_H5_NUM_NODES_RANK];
    const long long eoff = I[SOD_H5_RANK_ELEM_START] - 1, noff = I[SOD_H5_RANK_NODE_START] - 1;
    const int nnode = (int)I[SOD_H5_NNODE];
    std::vector<int> buf;
    int* conn_out[2] = {connecParOrig, connecParWork};
    const char* conn_name[2] = {"/Connectivity/connecParOrig", "/Connectivity/connecParWork"};
    for (int c = 0; c < 2; ++c) {
        if (conn_out[c] == nullptr) continue;
        buf.resize((size_t)(ne * nnode));
        if (int e = read_rows(f, conn_name[c], H5T_NATIVE_INT, eoff * nnode, ne * nnode, buf.data())) return e;
        elem_major_to_fortran(buf, ne, nnode, conn_out[c]);
    }
    if (workingNodesPar) {
        if (int e = read_rows(f, "/Connectivity/workingNodesPar", H5T_NATIVE_INT, m->work_off,
                              I[SOD_H5_NUM_WORKING_NODES], workingNodesPar))
            return e;
    }
    if (coordPar && nn > 0) {
        // read through FP32 and widened, as the reference's loader does
        std::vector<float> xyz((size_t)(nn * 3));
        if (int e = slab(f, m->coords.c_str(), H5T_NATIVE_FLOAT, (hsize_t)noff, (hsize_t)nn, 3, xyz.data(), false))
            return e;
        for (long long n = 0; n < nn; ++n)
            for (int d = 0; d < 3; ++d) coordPar[n + nn * d] = (double)xyz[n * 3 + d];
    }
    if (masSlaRankPar && I[SOD_H5_IS_PERIODIC]) {
        const long long np = I[SOD_H5_NPER];
        if (int e = read_rows(f, "/Periodic_data/masSlaRankPar1", H5T_NATIVE_INT, m->per_off, np, masSlaRankPar))
            return e;
        if (int e = read_rows(f, "/Periodic_data/masSlaRankPar2", H5T_NATIVE_INT, m->per_off, np, masSlaRankPar + np))
            return e;
    }
    if (globalIdSrl) {
        if (int e = read_rows(f, "/globalIds/globalIdSrl", H5T_NATIVE_LLONG, noff, nn, globalIdSrl)) return e;
    }
    if (globalIdPar) {
        if (int e = read_rows(f, "/globalIds/globalIdPar", H5T_NATIVE_LLONG, noff, nn, globalIdPar)) return e;
    }
    if (elemGid) {
        if (int e = read_rows(f, "/globalIds/elemGid", H5T_NATIVE_INT, eoff, ne, elemGid)) return e;
    }
    int* tab_out[3] = {a2ijk, gmsh2ijk, vtk2ijk};
    const char* tab_name[3] = {"/order/a2ijk", "/order/gmsh2ijk", "/order/vtk2ijk"};
    for (int t = 0; t < 3; ++t) {
        if (tab_out[t] == nullptr) continue;
        if (int e = read_rows(f, tab_name[t], H5T_NATIVE_INT, 0, nnode, tab_out[t])) return e;
    }
    return SOD_IO_OK;
}

int sod_h5_mesh_read_comms(const sod_h5_mesh* m, int* ranksToComm, int* commsMemPosInLoc, int* commsMemSize,
                           int* commsMemPosInNgb, int* nodesToComm) {
    if (m == nullptr) return fail(SOD_IO_ERR_ARG, "mesh_read_comms: NULL handle");
    if (m->info[SOD_H5_NRANKS] < 2) return SOD_IO_OK;
    const hid_t f = m->file;
    const long long nr = m->info[SOD_H5_NUM_RANKS_WITH_COMMS], nc = m->info[SOD_H5_NUM_NODES_TO_COMM];
    int* outs[4] = {ranksToComm, commsMemPosInLoc, commsMemSize, commsMemPosInNgb};
    const char* names[4] = {"/Parallel_data/ranksToComm", "/Parallel_data/commsMemPosInLoc",
                            "/Parallel_data/commsMemSize", "/Parallel_data/commsMemPosInNgb"};
    for (int t = 0; t < 4; ++t) {
        if (outs[t] == nullptr) continue;
        if (int e = read_rows(f, names[t], H5T_NATIVE_INT, m->rcomm_off, nr, outs[t])) return e;
    }
    if (nodesToComm) {
        if (int e = read_rows(f, "/Parallel_data/nodesToComm", H5T_NATIVE_INT, m->ncomm_off, nc, nodesToComm))
            return e;
    }
    return SOD_IO_OK;
}

void sod_h5_mesh_close(sod_h5_mesh* m) {
    if (m == nullptr) return;
    if (m->file >= 0) H5Fclose(m->file);
    delete m;
}

int sod_h5_mesh_create(const char* path, int nranks, int porder, const int* nelem_rank,
                       const long long* nnode_rank, const int* nwork_rank, const int* nper_rank,
                       const int* ncommranks_rank, const int* ncommnodes_rank, const int* a2ijk,
                       const int* gmsh2ijk, const int* vtk2ijk, const int* a2ij, const int* gmsh2ij,
                       const int* vtk2ij, sod_h5_mesh_writer** out) {
    if (out == nullptr || path == nullptr || nranks < 1 || nelem_rank == nullptr || nnode_rank == nullptr ||
        nwork_rank == nullptr || a2ijk == nullptr || gmsh2ijk == nullptr || vtk2ijk == nullptr ||
        a2ij == nullptr || gmsh2ij == nullptr || vtk2ij == nullptr)
        return fail(SOD_IO_ERR_ARG, "mesh_create: NULL or bad argument");
    if (nranks >= 2 && (ncommranks_rank == nullptr || ncommnodes_rank == nullptr))
        return fail(SOD_IO_ERR_ARG, "mesh_create: a partitioned file needs its communication list sizes");
    if (porder != 3) return fail(SOD_IO_ERR_PORDER, "mesh_create: this build writes p = 3 meshes only");
    *out = nullptr;
    quiet();
    const hid_t f = H5Fcreate(path, H5F_ACC_TRUNC, H5P_DEFAULT, H5P_DEFAULT);
    if (f < 0) return fail(SOD_IO_ERR_FILE, "cannot create mesh file %s", path);
    sod_h5_mesh_writer* w = new sod_h5_mesh_writer();
    w->file = f;
    w->nranks = nranks;
    w->periodic = nper_rank != nullptr;
    const int R = nranks, nnode = 64, npbou = 16;
    w->ne.assign(nelem_rank, nelem_rank + R);
    w->nn.assign(nnode_rank, nnode_rank + R);
    w->nw.assign(nwork_rank, nwork_rank + R);
    w->nper.assign(R, 0);
    w->ncr.assign(R, 0);
    w->ncn.assign(R, 0);
    if (nper_rank) w->nper.assign(nper_rank, nper_rank + R);
    if (R >= 2) {
        w->ncr.assign(ncommranks_rank, ncommranks_rank + R);
        w->ncn.assign(ncommnodes_rank, ncommnodes_rank + R);
    }
    w->done.assign(R, 0);
    auto bail = [&](int code) {
        H5Fclose(w->file);
        delete w;
        return code;
    };
    for (int r = 0; r < R; ++r)
        if (w->ne[r] < 0 || w->nn[r] < 0 || w->nw[r] < 0 || w->nper[r] < 0 || w->ncr[r] < 0 || w->ncn[r] < 0)
            return bail(fail(SOD_IO_ERR_ARG, "mesh_create: negative size for partition %d", r));
    const long long ntot = w->pre(w->nn, R), etot = w->pre(w->ne, R);
    const hid_t I4 = H5T_NATIVE_INT, I8 = H5T_STD_I64LE;
    int e = 0;
    // create_hdf5_groups_datasets_in_meshFile_from_tool (mod_hdf5.f90:176-306)
    const char* groups[] = {"/globalIds", "/Coords", "/Connectivity", "/order", "/meshOutputInfo"};
    for (const char* g : groups)
        if ((e = create_group(f, g))) return bail(e);
    if ((e = create1(f, "/globalIds/globalIdSrl", I8, ntot)) || (e = create1(f, "/globalIds/globalIdPar", I8, ntot)) ||
        (e = create1(f, "/globalIds/elemGid", I4, etot)))
        return bail(e);
    const hsize_t pdims[2] = {(hsize_t)ntot, 3};
    if ((e = create_dataset(f, "/Coords/Points", H5T_NATIVE_DOUBLE, 2, pdims))) return bail(e);
    if ((e = create1(f, "/Connectivity/connecParOrig", I4, etot * nnode)) ||
        (e = create1(f, "/Connectivity/connecParWork", I4, etot * nnode)) ||
        (e = create1(f, "/Connectivity/numWorkingNodesRankPar", I4, R)) ||
        (e = create1(f, "/Connectivity/workingNodesPar", I4, w->pre(w->nw, R))))
        return bail(e);
    if (R >= 2) {
        if ((e = create_group(f, "/Parallel_data"))) return bail(e);
        const char* scal[] = {"rankNodeStart", "rankNodeEnd", "rankElemStart", "rankElemEnd", "numRanksWithComms",
                              "numNodesToComm"};
        for (const char* s : scal)
            if ((e = create1(f, (std::string("/Parallel_data/") + s).c_str(), I4, R))) return bail(e);
        const char* lists[] = {"ranksToComm", "commsMemPosInLoc", "commsMemSize", "commsMemPosInNgb"};
        for (const char* s : lists)
            if ((e = create1(f, (std::string("/Parallel_data/") + s).c_str(), I4, w->pre(w->ncr, R)))) return bail(e);
        if ((e = create1(f, "/Parallel_data/nodesToComm", I4, w->pre(w->ncn, R)))) return bail(e);
    }
    if (w->periodic) {
        if ((e = create_group(f, "/Periodic_data")) || (e = create1(f, "/Periodic_data/nPerRankPar", I4, R)) ||
            (e = create1(f, "/Periodic_data/masSlaRankPar1", I4, w->pre(w->nper, R))) ||
            (e = create1(f, "/Periodic_data/masSlaRankPar2", I4, w->pre(w->nper, R))))
            return bail(e);
    }
    if ((e = create1(f, "/order/porder", I4, 1)) || (e = create1(f, "/order/a2ijk", I4, nnode)) ||
        (e = create1(f, "/order/gmsh2ijk", I4, nnode)) || (e = create1(f, "/order/vtk2ijk", I4, nnode)) ||
        (e = create1(f, "/order/a2ij", I4, npbou)) || (e = create1(f, "/order/gmsh2ij", I4, npbou)) ||
        (e = create1(f, "/order/vtk2ij", I4, npbou)))
        return bail(e);
    if ((e = create1(f, "/meshOutputInfo/isLinealOutput", H5T_STD_U8LE, 1)) ||
        (e = create1(f, "/meshOutputInfo/nnodeVTK", I4, 1)) ||
        (e = create1(f, "/meshOutputInfo/numVTKElemsPerMshElem", I4, 1)) ||
        (e = create1(f, "/meshOutputInfo/numElemsVTKMshRank", I4, R)) ||
        (e = create1(f, "/meshOutputInfo/sizeConnecVTKMshRank", I4, R)))
        return bail(e);
    // values that are not per-node or per-element
    if ((e = write_rows(f, "/order/porder", I4, 0, 1, &porder)) || (e = write_rows(f, "/order/a2ijk", I4, 0, nnode, a2ijk)) ||
        (e = write_rows(f, "/order/gmsh2ijk", I4, 0, nnode, gmsh2ijk)) ||
        (e = write_rows(f, "/order/vtk2ijk", I4, 0, nnode, vtk2ijk)) ||
        (e = write_rows(f, "/order/a2ij", I4, 0, npbou, a2ij)) || (e = write_rows(f, "/order/gmsh2ij", I4, 0, npbou, gmsh2ij)) ||
        (e = write_rows(f, "/order/vtk2ij", I4, 0, npbou, vtk2ij)))
        return bail(e);
    // high-order output: one VTK cell of nnode points per element
    // (mod_meshConversorTool.f90:254-260)
    const unsigned char lin = 0;
    const int nnode_vtk = nnode, per_elem = 1;
    std::vector<int> ev(R), cv(R);
    for (int r = 0; r < R; ++r) {
        ev[r] = w->ne[r] * per_elem;
        cv[r] = ev[r] * nnode_vtk;
    }
    if ((e = write_rows(f, "/meshOutputInfo/isLinealOutput", H5T_NATIVE_UCHAR, 0, 1, &lin)) ||
        (e = write_rows(f, "/meshOutputInfo/nnodeVTK", I4, 0, 1, &nnode_vtk)) ||
        (e = write_rows(f, "/meshOutputInfo/numVTKElemsPerMshElem", I4, 0, 1, &per_elem)) ||
        (e = write_rows(f, "/meshOutputInfo/numElemsVTKMshRank", I4, 0, R, ev.data())) ||
        (e = write_rows(f, "/meshOutputInfo/sizeConnecVTKMshRank", I4, 0, R, cv.data())) ||
        (e = write_rows(f, "/Connectivity/numWorkingNodesRankPar", I4, 0, R, w->nw.data())))
        return bail(e);
    if (w->periodic && (e = write_rows(f, "/Periodic_data/nPerRankPar", I4, 0, R, w->nper.data()))) return bail(e);
    if (R >= 2) {
        std::vector<int> ns(R), nend(R), es(R), eend(R);
        for (int r = 0; r < R; ++r) {
            ns[r] = (int)(w->pre(w->nn, r) + 1);
            nend[r] = (int)(w->pre(w->nn, r) + w->nn[r]);
            es[r] = (int)(w->pre(w->ne, r) + 1);
            eend[r] = (int)(w->pre(w->ne, r) + w->ne[r]);
        }
        if ((e = write_rows(f, "/Parallel_data/rankNodeStart", I4, 0, R, ns.data())) ||
            (e = write_rows(f, "/Parallel_data/rankNodeEnd", I4, 0, R, nend.data())) ||
            (e = write_rows(f, "/Parallel_data/rankElemStart", I4, 0, R, es.data())) ||
            (e = write_rows(f, "/Parallel_data/rankElemEnd", I4, 0, R, eend.data())) ||
            (e = write_rows(f, "/Parallel_data/numRanksWithComms", I4, 0, R, w->ncr.data())) ||
            (e = write_rows(f, "/Parallel_data/numNodesToComm", I4, 0, R, w->ncn.data())))
            return bail(e);
    }
    *out = w;
    return SOD_IO_OK;
}

int sod_h5_mesh_write_rank(sod_h5_mesh_writer* w, int rank, const int* connecParOrig, const int* connecParWork,
                           const int* workingNodesPar, const double* coordPar, const int* masSlaRankPar,
                           const long long* globalIdSrl, const long long* globalIdPar, const int* elemGid,
                           const int* ranksToComm, const int* commsMemPosInLoc, const int* commsMemSize,
                           const int* commsMemPosInNgb, const int* nodesToComm) {
    if (w == nullptr || rank < 0 || rank >= w->nranks) return fail(SOD_IO_ERR_ARG, "mesh_write_rank: bad handle or rank");
    const long long ne = w->ne[rank], nn = w->nn[rank], nw = w->nw[rank], np = w->nper[rank];
    if ((ne > 0 && (connecParOrig == nullptr || connecParWork == nullptr || elemGid == nullptr)) ||
        (nn > 0 && (coordPar == nullptr || globalIdSrl == nullptr || globalIdPar == nullptr)) ||
        (nw > 0 && workingNodesPar == nullptr) || (np > 0 && masSlaRankPar == nullptr))
        return fail(SOD_IO_ERR_ARG, "mesh_write_rank: a required array of partition %d is NULL", rank);
    const hid_t f = w->file, I4 = H5T_NATIVE_INT;
    const int nnode = 64;
    const long long eoff = w->pre(w->ne, rank), noff = w->pre(w->nn, rank);
    std::vector<int> buf;
    int e = 0;
    fortran_to_elem_major(connecParOrig, ne, nnode, buf);
    if ((e = write_rows(f, "/Connectivity/connecParOrig", I4, eoff * nnode, ne * nnode, buf.data()))) return e;
    fortran_to_elem_major(connecParWork, ne, nnode, buf);
    if ((e = write_rows(f, "/Connectivity/connecParWork", I4, eoff * nnode, ne * nnode, buf.data()))) return e;
    if ((e = write_rows(f, "/Connectivity/workingNodesPar", I4, w->pre(w->nw, rank), nw, workingNodesPar))) return e;
    if ((e = write_rows(f, "/globalIds/globalIdSrl", H5T_NATIVE_LLONG, noff, nn, globalIdSrl)) ||
        (e = write_rows(f, "/globalIds/globalIdPar", H5T_NATIVE_LLONG, noff, nn, globalIdPar)) ||
        (e = write_rows(f, "/globalIds/elemGid", I4, eoff, ne, elemGid)))
        return e;
    if (nn > 0) {
        std::vector<double> xyz((size_t)(nn * 3));
        for (long long n = 0; n < nn; ++n)
            for (int d = 0; d < 3; ++d) xyz[n * 3 + d] = coordPar[n + nn * d];
        if ((e = slab(f, "/Coords/Points", H5T_NATIVE_DOUBLE, (hsize_t)noff, (hsize_t)nn, 3, xyz.data(), true)))
            return e;
    }
    if (w->periodic) {
        const long long po = w->pre(w->nper, rank);
        if ((e = write_rows(f, "/Periodic_data/masSlaRankPar1", I4, po, np, masSlaRankPar)) ||
            (e = write_rows(f, "/Periodic_data/masSlaRankPar2", I4, po, np, masSlaRankPar + np)))
            return e;
    }
    if (w->nranks >= 2) {
        const long long nr = w->ncr[rank], nc = w->ncn[rank];
        if ((nr > 0 && (!ranksToComm || !commsMemPosInLoc || !commsMemSize || !commsMemPosInNgb)) ||
            (nc > 0 && !nodesToComm))
            return fail(SOD_IO_ERR_ARG, "mesh_write_rank: communication lists of partition %d are NULL", rank);
        const long long ro = w->pre(w->ncr, rank), co = w->pre(w->ncn, rank);
        if ((e = write_rows(f, "/Parallel_data/ranksToComm", I4, ro, nr, ranksToComm)) ||
            (e = write_rows(f, "/Parallel_data/commsMemPosInLoc", I4, ro, nr, commsMemPosInLoc)) ||
            (e = write_rows(f, "/Parallel_data/commsMemSize", I4, ro, nr, commsMemSize)) ||
            (e = write_rows(f, "/Parallel_data/commsMemPosInNgb", I4, ro, nr, commsMemPosInNgb)) ||
            (e = write_rows(f, "/Parallel_data/nodesToComm", I4, co, nc, nodesToComm)))
            return e;
    }
    w->done[rank] = 1;
    return SOD_IO_OK;
}

int sod_h5_mesh_writer_close(sod_h5_mesh_writer* w) {
    if (w == nullptr) return SOD_IO_OK;
    int missing = -1;
    for (int r = 0; r < w->nranks; ++r)
        if (!w->done[r]) missing = r;
    const herr_t st = H5Fclose(w->file);
    delete w;
    if (missing >= 0) return fail(SOD_IO_ERR_ARG, "mesh_writer_close: partition %d was never written", missing);
    if (st < 0) return fail(SOD_IO_ERR_FILE, "mesh_writer_close: closing the file failed");
    return SOD_IO_OK;
}

}  // extern "C"

// ---------------------------------------------------------------- restart

namespace {

const char* kNodeFields[] = {"rho", "u_x", "u_y", "u_z", "pr", "E", "mue", "mut"};
const char* kWalave[] = {"walave_u_x", "walave_u_y", "walave_u_z"};

int check_conn(const int* conn, long long n, int npoin) {
    for (long long t = 0; t < n; ++t)
        if (conn[t] < 1 || conn[t] > npoin) return fail(SOD_IO_ERR_ARG, "connecParOrig holds an id outside 1..npoin");
    return SOD_IO_OK;
}

int write_f32(hid_t f, const char* name, long long off, int n, const double* v) {
    std::vector<float> t((size_t)n);
    for (int i = 0; i < n; ++i) t[i] = (float)v[i];   // real(x, rp_vtk)
    return write_rows(f, name, H5T_NATIVE_FLOAT, off, n, t.data());
}

int read_f32(hid_t f, const char* name, long long off, int n, double* v) {
    std::vector<float> t((size_t)n);
    if (int e = read_rows(f, name, H5T_NATIVE_FLOAT, off, n, t.data())) return e;
    for (int i = 0; i < n; ++i) v[i] = (double)t[i];
    return SOD_IO_OK;
}

}  // namespace

extern "C" {

int sod_h5_restart_save(const char* path, int create, long long total_nodes, long long node_start, int npoin,
                        int nelem, const int* connecParOrig, int istep, double time, const double* rho,
                        const double* u, const double* pr, const double* E, const double* mu_e,
                        const double* mu_t, const double* walave_u) {
    if (path == nullptr || npoin < 0 || nelem < 0 || node_start < 1 || node_start - 1 + npoin > total_nodes ||
        (npoin > 0 && (!rho || !u || !pr || !E)) || (nelem > 0 && (!connecParOrig || !mu_e || !mu_t)))
        return fail(SOD_IO_ERR_ARG, "restart_save: bad argument");
    if (int e = check_conn(connecParOrig, (long long)nelem * 64, npoin)) return e;
    quiet();
    const hid_t f = create ? H5Fcreate(path, H5F_ACC_TRUNC, H5P_DEFAULT, H5P_DEFAULT)
                           : H5Fopen(path, H5F_ACC_RDWR, H5P_DEFAULT);
    if (f < 0) return fail(SOD_IO_ERR_FILE, "cannot %s restart file %s", create ? "create" : "open", path);
    H5Id file(f, H5Fclose);
    int e = 0;
    if (create) {
        for (const char* nm : kNodeFields)
            if ((e = create1(f, nm, H5T_NATIVE_FLOAT, total_nodes))) return e;
        if (walave_u)
            for (const char* nm : kWalave)
                if ((e = create1(f, nm, H5T_NATIVE_FLOAT, total_nodes))) return e;
        const float tf = (float)time;
        if ((e = create1(f, "time", H5T_NATIVE_FLOAT, 1)) || (e = write_rows(f, "time", H5T_NATIVE_FLOAT, 0, 1, &tf)) ||
            (e = create1(f, "istep", H5T_NATIVE_INT, 1)) || (e = write_rows(f, "istep", H5T_NATIVE_INT, 0, 1, &istep)))
            return e;
    }
    // element-node viscosities -> node arrays (mod_hdf5.f90:3235-3241)
    std::vector<double> aux_e((size_t)npoin, 0.0), aux_t((size_t)npoin, 0.0);
    for (int el = 0; el < nelem; ++el)
        for (int a = 0; a < 64; ++a) {
            const size_t ea = (size_t)el + (size_t)nelem * a;
            const int n = connecParOrig[ea] - 1;
            aux_e[n] = mu_e[ea];
            aux_t[n] = mu_t[ea];
        }
    const long long off = node_start - 1;
    const size_t np = (size_t)npoin;
    const double* src[8] = {rho, u, u + np, u + 2 * np, pr, E, aux_e.data(), aux_t.data()};
    for (int k = 0; k < 8; ++k)
        if ((e = write_f32(f, kNodeFields[k], off, npoin, src[k]))) return e;
    if (walave_u)
        for (int d = 0; d < 3; ++d)
            if ((e = write_f32(f, kWalave[d], off, npoin, walave_u + d * np))) return e;
    return SOD_IO_OK;
}

int sod_h5_restart_load(const char* path, long long node_start, int npoin, int nelem,
                        const int* connecParOrig, int nper, const int* masSlaRankPar, int* istep, double* time,
                        double* rho, double* u, double* pr, double* E, double* mu_e, double* mu_t,
                        double* walave_u) {
    if (path == nullptr || npoin < 0 || nelem < 0 || nper < 0 || node_start < 1 || !istep || !time ||
        (npoin > 0 && (!rho || !u || !pr || !E)) || (nelem > 0 && (!connecParOrig || !mu_e || !mu_t)) ||
        (nper > 0 && !masSlaRankPar))
        return fail(SOD_IO_ERR_ARG, "restart_load: bad argument");
    if (int e = check_conn(connecParOrig, (long long)nelem * 64, npoin)) return e;
    for (long long t = 0; t < 2LL * nper; ++t)
        if (masSlaRankPar[t] < 1 || masSlaRankPar[t] > npoin)
            return fail(SOD_IO_ERR_ARG, "masSlaRankPar holds an id outside 1..npoin");
    quiet();
    const hid_t f = H5Fopen(path, H5F_ACC_RDONLY, H5P_DEFAULT);
    if (f < 0) return fail(SOD_IO_ERR_FILE, "cannot open restart file %s", path);
    H5Id file(f, H5Fclose);
    int e = 0;
    float tf = 0.0f;
    if ((e = read_rows(f, "time", H5T_NATIVE_FLOAT, 0, 1, &tf)) || (e = read_rows(f, "istep", H5T_NATIVE_INT, 0, 1, istep)))
        return e;
    *time = (double)tf;
    const long long off = node_start - 1;
    const size_t np = (size_t)npoin;
    std::vector<double> aux_e(np), aux_t(np);
    double* dst[8] = {rho, u, u + np, u + 2 * np, pr, E, aux_e.data(), aux_t.data()};
    for (int k = 0; k < 8; ++k)
        if ((e = read_f32(f, kNodeFields[k], off, npoin, dst[k]))) return e;
    if (walave_u) {
        if (exists(f, kWalave[0])) {
            for (int d = 0; d < 3; ++d)
                if ((e = read_f32(f, kWalave[d], off, npoin, walave_u + d * np))) return e;
        } else {
            std::memcpy(walave_u, u, sizeof(double) * 3 * np);
        }
    }
    for (int el = 0; el < nelem; ++el)
        for (int a = 0; a < 64; ++a) {
            const size_t ea = (size_t)el + (size_t)nelem * a;
            const int n = connecParOrig[ea] - 1;
            mu_e[ea] = aux_e[n];
            mu_t[ea] = aux_t[n];
        }
    // periodic slaves take their master's values (mod_hdf5.f90:3409-3430)
    for (int i = 0; i < nper; ++i) {
        const size_t ms = (size_t)masSlaRankPar[i] - 1, sl = (size_t)masSlaRankPar[nper + i] - 1;
        rho[sl] = rho[ms];
        for (int d = 0; d < 3; ++d) u[sl + d * np] = u[ms + d * np];
        pr[sl] = pr[ms];
        E[sl] = E[ms];
        if (walave_u)
            for (int d = 0; d < 3; ++d) walave_u[sl + d * np] = walave_u[ms + d * np];
    }
    return SOD_IO_OK;
}

}  // extern "C"
