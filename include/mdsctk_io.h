/*
 * mdsctk_io.h - host file formats of the knn_rms path.
 *
 * Replaces the GROMACS file I/O the reference calls (knn_rms.cpp:149-206):
 *   read_tps_conf            -> mdsctk_topology_read (PDB/GRO, masses)
 *   open_xtc/read_first_xtc/read_next_xtc -> mdsctk_xtc_info / mdsctk_xtc_read
 * and provides a writer (mdsctk_xtc_write) for fixtures and synthetic data.
 * All functions return >= 0 on success or a negative MDSCTK_ERR_* code
 * (mdsctk_knn.h); mdsctk_io_last_error() describes the last failure of the
 * calling thread.
 */
#ifndef MDSCTK_IO_H
#define MDSCTK_IO_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Count frames and read the atom count and first-frame precision of an XTC
 * file (walks the frame headers; coordinates are not decoded). */
int mdsctk_xtc_info(const char* path, int64_t* n_frames, int32_t* natoms, float* precision);

/* Decode up to max_frames frames into xyz[n][natoms][3] (nm).  natoms must
 * match the file (MDSCTK_ERR_NATOMS otherwise, as knn_rms.cpp:158-178).
 * steps/times/boxes (int[n], float[n], float[n][9]) may be NULL.  Frames are
 * decoded on `nthreads` host threads (<= 0: all cores).  *n_read receives the
 * number of frames decoded. */
int mdsctk_xtc_read(const char* path, float* xyz, int64_t max_frames, int32_t natoms,
                    int64_t* n_read, int32_t* steps, float* times, float* boxes, int nthreads);

/* Decode a whole XTC file on HIP device `device`: the file is read into
 * pinned host chunks, each uploaded while the next is read, the frame
 * headers are walked as the chunks arrive, and the frames are decoded on
 * the device, one lane per frame (same values as mdsctk_xtc_read).
 * *d_xyz receives a device buffer [n][natoms][3] (nm) that the caller
 * releases with mdsctk_xtc_free_device (NULL for an empty file), *n_read the
 * frame count.  natoms must match the file (MDSCTK_ERR_NATOMS otherwise, as
 * knn_rms.cpp:158-178).  Replaces the host-side read loop of
 * knn_rms.cpp:186-206 for callers that compute on the device. */
int mdsctk_xtc_read_device(const char* path, int32_t natoms, int device, float** d_xyz,
                           int64_t* n_read);

/* Atom count of the first frame of an XTC file (reads its header only). */
int mdsctk_xtc_natoms(const char* path, int32_t* natoms);

/* Release a buffer returned by mdsctk_xtc_read_device. */
int mdsctk_xtc_free_device(float* d_xyz);

/* Write frames as XTC with the given precision (e.g. 1000 = 1e-3 nm).
 * steps/times/boxes may be NULL (step = frame index, time = 0, zero box). */
int mdsctk_xtc_write(const char* path, const float* xyz, int64_t n_frames, int32_t natoms,
                     float precision, const int32_t* steps, const float* times,
                     const float* boxes);

/* Read a PDB (.pdb/.ent, first model) or GRO (.gro) topology: atom count and
 * per-atom masses.  Masses come from the element column when present,
 * otherwise from the atom name (see DESIGN notes in the source).  Pass
 * masses == NULL to query natoms only. */
int mdsctk_topology_read(const char* path, int32_t* natoms, float* masses, int32_t capacity);

/* Message of the last I/O failure on this thread (never NULL). */
const char* mdsctk_io_last_error(void);

#ifdef __cplusplus
}
#endif

#endif /* MDSCTK_IO_H */
