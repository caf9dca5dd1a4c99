/*
 * nlod.h — C ABI of the MI355X (gfx950) neural-LoD hot path.
 *
 * Every function returns an int status (NLOD_OK == 0); on failure the message
 * of the calling thread's last error is available from nlod_last_error().
 * No C++ exception crosses this boundary.  Device buffers passed in are owned
 * by the caller; model and trainer handles own their weights and workspaces.
 * Calls taking a hipStream_t (passed as void*) are asynchronous and ordered on
 * that stream unless documented otherwise.  Handles are not thread-safe.
 *
 * Each entry point names the reference interface it replaces (paths relative
 * to the neural_lod repository).
 */
#ifndef NLOD_H
#define NLOD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ------------------------------------------------------------------ status */

enum nlod_status {
    NLOD_OK = 0,
    NLOD_ERR_INVALID_ARGUMENT = 1, /* bad pointer, size, shape or flag (tcnn CHECK_THROW) */
    NLOD_ERR_CONFIG = 2,           /* malformed or unsupported JSON configuration */
    NLOD_ERR_UNSUPPORTED = 3,      /* valid request this build does not implement */
    NLOD_ERR_HIP = 4,              /* a HIP runtime call failed */
    NLOD_ERR_OUT_OF_MEMORY = 5,
    NLOD_ERR_STATE = 6             /* call out of order (e.g. backward without forward) */
};

/* Message of the calling thread's most recent failure ("" if none).
 * Replaces the std::runtime_error texts of CHECK_THROW / checkCuda
 * (rptr/cuda/render_cuda.h:8-16, tiny-cuda-nn/common.h). */
const char* nlod_last_error(void);

/* Library version string. */
const char* nlod_version(void);

/* -------------------------------------------------------------- voxelizer */

/* Voxelization parameters.  Mirrors struct voxinfo
 * (neural_lod/cuda_voxelizer/src/util.h:57-76). */
typedef struct nlod_voxinfo {
    float bbox_min[3];
    float bbox_max[3];
    uint32_t gridsize[3];
    uint32_t reserved0;
    uint64_t n_triangles;
    float unit[3]; /* (bbox_max - bbox_min) / gridsize, per axis, in float */
    uint32_t reserved1;
} nlod_voxinfo;

enum nlod_vox_flags {
    NLOD_VOX_SURFACE = 0,      /* conservative surface voxelization (voxelize.cu:58-190) */
    NLOD_VOX_SOLID = 1u << 0,  /* parity solid fill (voxelize_solid.cu:73-146, -solid) */
    NLOD_VOX_MORTON = 1u << 1, /* Morton-ordered bits (-o morton); needs a power-of-two grid */
    NLOD_VOX_PAD_EPS = 1u << 2,/* pad the cube by 1e-4 of its extent (main.cpp:45,209-215, -pad) */
    NLOD_VOX_ACCUMULATE = 1u << 3 /* do not clear the table first */
};

/* Number of uint32 words of a res^3 bit table: ceil(res^3 / 32)
 * (main.cpp:220). */
int nlod_vox_table_words(uint32_t res, uint64_t* words_out);

/* Cubic voxelization box from a mesh box, optionally padded by 1e-4 of its
 * extent.  Replaces createMeshBBCube (util.h:87-112) + the voxinfo
 * constructor (util.h:63-68) as called at main.cpp:214-216.  Host only. */
int nlod_voxinfo_from_bbox(const float bbox_min[3], const float bbox_max[3], uint32_t res,
                           uint32_t flags, uint64_t n_triangles, nlod_voxinfo* info_out);

/* Per-vertex bounding box of a device triangle soup float[n][9] (exact
 * min/max; replaces trimesh2 need_bbox, TriMesh_bounding.cc:18-31).
 * Synchronizes the stream to return the result to the host. */
int nlod_mesh_bbox(void* stream, const float* d_triangles, uint64_t n_triangles,
                   float bbox_min_out[3], float bbox_max_out[3]);

/* Voxelize a device triangle soup float[n][9] (x0 y0 z0 x1 ... z2 per
 * triangle) into the device table d_table of nlod_vox_table_words(G) words.
 * Clears the table first unless NLOD_VOX_ACCUMULATE.  Bit 31-(i%32) of word
 * i/32 holds voxel i = x + y*G + z*G*G.  Results are bit-identical to the
 * reference CPU voxelizer (cpu_voxelizer.cpp:39-193 / :252-328).
 * Replaces void voxelize(const voxinfo&, float*, unsigned*, bool, bool)
 * (voxelize.cu:192) and voxelize_solid(...) (voxelize_solid.cu:148).
 * Asynchronous. */
int nlod_voxelize(void* stream, const nlod_voxinfo* info, const float* d_triangles,
                  uint32_t* d_table, uint32_t flags);

/* voxelize(mesh, res): bounding box, cube, then nlod_voxelize.  Writes the
 * voxinfo used to info_out (may be NULL).  Synchronizes once (bounding box),
 * the voxelization itself is asynchronous.  Replaces the GPU branch of
 * main.cpp:199-258. */
int nlod_voxelize_mesh(void* stream, const float* d_triangles, uint64_t n_triangles, uint32_t res,
                       uint32_t flags, uint32_t* d_table, nlod_voxinfo* info_out);

#ifdef __cplusplus
}
#endif

#endif /* NLOD_H */
