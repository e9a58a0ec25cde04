/*
 * curvami_h5.h -- C ABI of the step container: the HDF5 file that
 * Exporter::writeBinaryStep (plate/Exporter.h:27-339) appends one step to per
 * call, written through the HDF5 C library with the reference's layout, and
 * read back by name. Host only.
 *
 * Layout per step (Exporter.h:203-322):
 *  - dataset /step_NNNNN: num_gps x num_cols doubles [X, Y, u, v, w, rotX_deg,
 *    rotY_deg, fields...], chunked (min(num_gps, 1000), num_cols), deflate 6;
 *  - dataset /step_NNNNN_creases (only when it has rows): rows x 7 doubles
 *    [fold_index, crease_tag, X_ref, Y_ref, s_arc, theta_deg, weight], same
 *    chunking and compression, with the string attribute "columns";
 *  - attributes of /step_NNNNN: git_commit (string), angles (array of doubles,
 *    one 0 when empty), is_driver (array of ints, one 0 when empty), disp,
 *    W_ext, E_mem, E_ben, E_hin, planar_rms, planar_max, ref_diag (doubles).
 * The file is created only when it does not exist and otherwise opened for
 * writing, so a failed open never truncates the steps already written
 * (Exporter.h:206-216); writing a step that exists fails.
 *
 * Status codes: 0 ok, 1 invalid argument, 2 HDF5 failure (message in
 * cvm_h5_last_error, per thread).
 */
#ifndef CURVAMI_H5_H
#define CURVAMI_H5_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define CVM_H5_API __attribute__((visibility("default")))

CVM_H5_API const char* cvm_h5_last_error(void);

/* scalars8: disp, W_ext, E_mem, E_ben, E_hin, planar_rms, planar_max, ref_diag. */
CVM_H5_API int cvm_h5_write_step(const char* path, int step, const double* data, int64_t num_gps, int num_cols,
                                 const double* creases, int64_t crease_rows, const double* angles, int nangles,
                                 const int* is_driver, int ndriver, const double* scalars8, const char* git_commit);

/* A 2-D double dataset by name: dims[2] always; values (row-major) when
 * out != NULL and cap >= dims[0] * dims[1]. */
CVM_H5_API int cvm_h5_read_dataset(const char* path, const char* name, int64_t* dims, double* out, int64_t cap);

/* A numeric attribute of a dataset as doubles (scalar or array, double or int):
 * *n values, written to out when n <= cap. */
CVM_H5_API int cvm_h5_read_attr(const char* path, const char* dataset, const char* attr, double* out, int cap,
                                int* n);

/* Names of the datasets in the file's root group, comma-separated, into buf
 * (NUL-terminated; *needed = the length the full list needs). */
CVM_H5_API int cvm_h5_list(const char* path, char* buf, int cap, int* needed);

/* A string attribute into buf (NUL-terminated, truncated to cap). */
CVM_H5_API int cvm_h5_read_attr_string(const char* path, const char* dataset, const char* attr, char* buf,
                                       int cap);

#ifdef __cplusplus
}
#endif
#endif /* CURVAMI_H5_H */
