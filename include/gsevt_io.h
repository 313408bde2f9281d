/* C ABI of libgsevt_io.so: host-side readers of the tracker's on-disk inputs.
 * Plain host pointers and sizes; every call returns 0 on success and -1 on
 * error, with the message in gsevt_io_last_error() (thread-local).  The
 * library allocates nothing the caller sees: output arrays are sized with the
 * *_info / *_lines queries first. */
#ifndef GSEVT_IO_H_
#define GSEVT_IO_H_

#ifdef __cplusplus
extern "C" {
#endif

const char* gsevt_io_last_error(void);

/* 3DGS PLY map, replacing GaussianModel.load_ply
 * (gaussian_splatting/scene/gaussian_model.py:238-327, via plyfile).
 * Sizes of the vertex element: vertex count, number of f_rest_*, scale_* and
 * rot_* properties. */
int gsevt_ply_info(const char* path, long long* n_vertices, int* n_rest, int* n_scale, int* n_rot);
/* Vertex properties as float32, rows per vertex: xyz [n][3], normals [n][3]
 * (nullable), f_dc [n][3] (f_dc_0..2), f_rest [n][n_rest] in f_rest index
 * order, opacity [n], scales [n][n_scale] and rots [n][n_rot] in index order.
 * ascii and binary little / big endian; any property order and scalar type. */
int gsevt_ply_read(const char* path, float* xyz, float* normals, float* f_dc, float* f_rest, float* opacity,
                   float* scales, float* rots);

/* events.txt stream, replacing the line loop of load_events_from_txt
 * (utils/event_camera/event.py:11-39).  Line count (upper bound for the
 * arrays), or -1. */
long long gsevt_events_txt_lines(const char* path);
/* Parses every line into ts / x / y / p (fields 0..3 of the line's integers,
 * split on single spaces as line.strip().split(' ') does) in file order;
 * *n_out = lines parsed.  *bad_line = index of the first line that does not
 * parse (its slots hold 0) or -1, with *bad_kind 1 (a field is not an
 * integer), 2 (fewer than 4 fields) or 3 (x, y or p outside int32); the
 * caller decides, as the reference does, whether that line is reached. */
int gsevt_events_txt_read(const char* path, long long capacity, long long* ts, int* x, int* y, int* p,
                          long long* n_out, long long* bad_line, int* bad_kind);

#ifdef __cplusplus
}  /* extern "C" */
#endif

#endif  /* GSEVT_IO_H_ */
