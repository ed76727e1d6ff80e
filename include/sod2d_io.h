/*
 * libsod2d_io: SOD2D's partitioned HDF5 mesh files and restart files, read
 * and written on the host through the HDF5 C library.
 *
 * These entry points stand in for the Fortran HDF5 layer of the reference
 * (src/lib_sod2d/sources/mod_hdf5.f90), so that a run on the device can
 * start from the reference's own mesh files and resume from, or leave, its
 * restart files.  Files hold the reference's layout exactly: the same group
 * and dataset names, shapes and element types.
 *
 * Conventions (as in libsod2d_hip):
 *   - every function returns 0 on success or a negative SOD_IO_ERR_* code;
 *     sod_io_last_error() gives the message of the last failure (per thread);
 *   - arrays are caller-owned, Fortran column-major, with the extents named
 *     in each comment; node and element ids are 1-based as in the files;
 *   - "rank" is a mesh partition (mshRank) of a file made for `nranks`
 *     partitions: its nodes, elements, working nodes, periodic pairs and
 *     communication lists are one contiguous slice of each dataset.
 */
#ifndef SOD2D_IO_H
#define SOD2D_IO_H

#ifdef __cplusplus
extern "C" {
#endif

enum {
    SOD_IO_OK = 0,
    SOD_IO_ERR_ARG = -1,      /* bad argument (NULL where required, size) */
    SOD_IO_ERR_FILE = -2,     /* file missing, unreadable or not writable */
    SOD_IO_ERR_FORMAT = -3,   /* dataset missing or of an unexpected shape */
    SOD_IO_ERR_PORDER = -4    /* mesh order other than p = 3 */
};

const char* sod_io_last_error(void);

/* ------------------------------------------------------------ file names */

/* set_hdf5_meshFile_name (mod_hdf5.f90:83-91):
 * <path><name>-<nranks>.hdf.  Returns the length written (without the
 * terminating NUL) or SOD_IO_ERR_ARG when `cap` is too small. */
int sod_h5_mesh_file_name(const char* path, const char* name, int nranks, char* out, int cap);

/* set_hdf5_baseResultsFile_name + set_hdf5_restartFile_name
 * (mod_hdf5.f90:99-114, 2836-2844): <path>restart_<mesh>-<nranks>_<cnt>.h5 */
int sod_h5_restart_file_name(const char* path, const char* mesh_name, int nranks, int restart_cnt, char* out,
                             int cap);

/* ---------------------------------------------------------- mesh reading */

typedef struct sod_h5_mesh sod_h5_mesh;

/* Indices into the info vector of sod_h5_mesh_info. */
enum {
    SOD_H5_NRANKS = 0,          /* partitions in the file (numWorkingNodesRankPar extent) */
    SOD_H5_PORDER,              /* /order/porder (always 3 after a successful open) */
    SOD_H5_NNODE,               /* nodes per element, (porder+1)^3 */
    SOD_H5_TOTAL_NODES,         /* totalNumNodesPar (/globalIds/globalIdSrl extent) */
    SOD_H5_TOTAL_ELEMS,         /* totalNumElements (/globalIds/elemGid extent) */
    SOD_H5_IS_PERIODIC,         /* /Periodic_data present */
    SOD_H5_IS_LINEAL_OUTPUT,    /* /meshOutputInfo/isLinealOutput */
    SOD_H5_HAS_BOUNDARIES,      /* /Boundary_data present (not read: the periodic path has none) */
    SOD_H5_RANK_NODE_START,     /* rankNodeStart, 1-based */
    SOD_H5_RANK_NODE_END,
    SOD_H5_RANK_ELEM_START,
    SOD_H5_RANK_ELEM_END,
    SOD_H5_NUM_NODES_RANK,      /* numNodesRankPar  (npoin of this partition) */
    SOD_H5_NUM_ELEMS_RANK,      /* numElemsRankPar  (nelem of this partition) */
    SOD_H5_NUM_WORKING_NODES,   /* numWorkingNodesRankPar (npoin_w) */
    SOD_H5_NPER,                /* nPerRankPar, periodic (master, slave) pairs */
    SOD_H5_NUM_RANKS_WITH_COMMS,
    SOD_H5_NUM_NODES_TO_COMM,
    SOD_H5_INFO_COUNT
};

/* Opens a mesh file and reads the sizes of partition `rank`
 * (load_hdf5_meshFile, mod_hdf5.f90:1294-1401: globalIds extents,
 * meshOutputInfo, porder, Parallel_data when nranks >= 2, Periodic_data,
 * Connectivity/numWorkingNodesRankPar).  A mesh of order other than 3 is
 * rejected with SOD_IO_ERR_PORDER, as the reference aborts
 * (mod_hdf5.f90:2779-2782). */
int sod_h5_mesh_open(const char* path, int rank, sod_h5_mesh** out);

/* info[SOD_H5_INFO_COUNT] of the open partition. */
int sod_h5_mesh_info(const sod_h5_mesh* m, long long* info);

/* The arrays of the open partition (any pointer may be NULL to skip it):
 *   connecParOrig, connecParWork  (nelem,nnode)  load_connectivity_hdf5, mod_hdf5.f90:2117-2165
 *   workingNodesPar (npoin_w)                    mod_hdf5.f90:2167-2202
 *   coordPar (npoin,3), rounded through FP32 as the reference reads it
 *       (read_array2D_tr_rp_in_dataset_hdf5_file with rp_vtk = 4,
 *       mod_hdf5.f90:3030-3051; load_coordinates_hdf5, :2648-2672)
 *   masSlaRankPar (nper,2)                       load_periodic_data_hdf5, :2404-2471
 *   globalIdSrl, globalIdPar (npoin), elemGid (nelem)   load_globalIds_hdf5, :2674-2715
 *   a2ijk, gmsh2ijk, vtk2ijk (nnode)             load_porder_hdf5, :2761-2819 */
int sod_h5_mesh_read(const sod_h5_mesh* m, int* connecParOrig, int* connecParWork, int* workingNodesPar,
                     double* coordPar, int* masSlaRankPar, long long* globalIdSrl, long long* globalIdPar,
                     int* elemGid, int* a2ijk, int* gmsh2ijk, int* vtk2ijk);

/* Communication lists of the open partition (load_parallel_data_hdf5,
 * mod_hdf5.f90:2206-2320): ranksToComm, commsMemPosInLoc, commsMemSize,
 * commsMemPosInNgb (numRanksWithComms), nodesToComm (numNodesToComm).
 * All zero-length for a one-partition file. */
int sod_h5_mesh_read_comms(const sod_h5_mesh* m, int* ranksToComm, int* commsMemPosInLoc, int* commsMemSize,
                           int* commsMemPosInNgb, int* nodesToComm);

void sod_h5_mesh_close(sod_h5_mesh* m);

/* ---------------------------------------------------------- mesh writing */

typedef struct sod_h5_mesh_writer sod_h5_mesh_writer;

/* Creates a mesh file for `nranks` partitions with every group and dataset
 * of the mesh tool's output (create_hdf5_groups_datasets_in_meshFile_from_tool,
 * mod_hdf5.f90:176-306, without the /VTKHDF visualisation group and without
 * boundary data).  Per-partition sizes: nelem_rank, nnode_rank, nwork_rank,
 * nper_rank (NULL: not periodic), ncommranks_rank and ncommnodes_rank (read
 * only when nranks >= 2).  Order tables: a2ijk, gmsh2ijk, vtk2ijk (64) and
 * a2ij, gmsh2ij, vtk2ij (16), written as given. */
int sod_h5_mesh_create(const char* path, int nranks, int porder, const int* nelem_rank,
                       const long long* nnode_rank, const int* nwork_rank, const int* nper_rank,
                       const int* ncommranks_rank, const int* ncommnodes_rank, const int* a2ijk,
                       const int* gmsh2ijk, const int* vtk2ijk, const int* a2ij, const int* gmsh2ij,
                       const int* vtk2ij, sod_h5_mesh_writer** out);

/* Writes partition `rank` (write_mshRank_data_in_hdf5_meshFile_from_tool,
 * mod_hdf5.f90:583-1060): same arrays and layouts as sod_h5_mesh_read plus
 * the communication lists (ignored when nranks == 1).  coordPar is written
 * in FP64 as given. */
int sod_h5_mesh_write_rank(sod_h5_mesh_writer* w, int rank, const int* connecParOrig, const int* connecParWork,
                           const int* workingNodesPar, const double* coordPar, const int* masSlaRankPar,
                           const long long* globalIdSrl, const long long* globalIdPar, const int* elemGid,
                           const int* ranksToComm, const int* commsMemPosInLoc, const int* commsMemSize,
                           const int* commsMemPosInNgb, const int* nodesToComm);

int sod_h5_mesh_writer_close(sod_h5_mesh_writer* w);

/* --------------------------------------------------------- restart files */

/* save_hdf5_restartFile (mod_hdf5.f90:3214-3312) for one partition: the
 * node slice [node_start, node_start+npoin) of datasets of total_nodes
 * entries.  rho, pr, E (npoin), u and walave_u (npoin,3; walave_u may be
 * NULL), mu_e and mu_t (nelem,nnode) element-node fields, which are stored
 * per node through connecParOrig (element loop outside, node loop inside,
 * so the last element that touches a node gives its value,
 * mod_hdf5.f90:3235-3241).  Fields and `time` are stored as FP32 (rp_vtk),
 * istep as int32.  create = 1 creates the file and its datasets (and writes
 * time and istep); create = 0 writes the slice into an existing file. */
int sod_h5_restart_save(const char* path, int create, long long total_nodes, long long node_start, int npoin,
                        int nelem, const int* connecParOrig, int istep, double time, const double* rho,
                        const double* u, const double* pr, const double* E, const double* mu_e,
                        const double* mu_t, const double* walave_u);

/* load_hdf5_restartFile (mod_hdf5.f90:3314-3436) for one partition: reads
 * the node slice through FP32, widens to FP64, scatters mue/mut back to the
 * element nodes through connecParOrig, and copies rho, u, pr, E (and
 * walave_u) from each periodic master to its slave (masSlaRankPar(nper,2),
 * 1-based local ids).  With walave_u non-NULL and no walave datasets in the
 * file, walave_u = u.  time is read through FP32 as the reference does. */
int sod_h5_restart_load(const char* path, long long node_start, int npoin, int nelem,
                        const int* connecParOrig, int nper, const int* masSlaRankPar, int* istep, double* time,
                        double* rho, double* u, double* pr, double* E, double* mu_e, double* mu_t,
                        double* walave_u);

#ifdef __cplusplus
}
#endif

#endif /* SOD2D_IO_H */
