/* C ABI of libgsevt_io.so: host-side readers of the tracker's on-disk inputs.
 * Plain host pointers and sizes.  Calls returning int return 0 on success and
 * -1 on error (gsevt_events_txt_lines: the count, or -1); gsevt_bag_open
 * returns NULL on error.  The message of an error is in gsevt_io_last_error()
 * (thread-local).  Output arrays are the caller's, sized with the *_info /
 * *_lines / *_open queries first. */
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

/* ROS1 bag (format 2.0) event stream, replacing the rosbag loop of
 * load_events_from_bag (utils/event_camera/event.py:42-69).  Messages of
 * `topic` are those rosbag's read_messages(topics=[topic]) yields: found
 * through the bag's index (an unindexed bag is an error), in bag time order
 * with ties in file order, and those whose bag time secs + nsecs / 1e9 (in
 * double, as Time.to_sec()) is below start_time dropped when has_start != 0.
 * The topic's connections must carry the wire layout of dvs_msgs/EventArray,
 * which prophesee_event_msgs/EventArray shares (checked on the stored message
 * definition).  Chunks: none, bz2 and lz4 (libbz2 / liblz4 loaded on first
 * use).
 *
 * gsevt_bag_open maps the file, decompresses the topic's chunks once and
 * returns an opaque handle (NULL on error) with the counts of kept messages
 * and of their events; the handle holds the uncompressed chunks until
 * gsevt_bag_close.  This handle is the one allocation of this library that
 * outlives a call. */
typedef struct gsevt_bag gsevt_bag;
gsevt_bag* gsevt_bag_open(const char* path, const char* topic, int has_start, double start_time,
                          long long* n_messages, long long* n_events);
/* The kept messages' events in order into arrays of n_events (msg_events: of
 * n_messages): x, y, p (1 where the polarity byte is nonzero, as
 * `1 if polarity else 0`) and ts_us, the message's header.stamp
 * (event_stamps == 0, what the reference stores) or the event's own ts
 * (event_stamps != 0), in integer microseconds (secs * 10^6 + nsecs / 1000).
 * msg_events[i] = events of message i. */
int gsevt_bag_events(const gsevt_bag* bag, int event_stamps, long long* msg_events, long long* ts_us, int* x, int* y,
                     int* p);
void gsevt_bag_close(gsevt_bag* bag);

#ifdef __cplusplus
}  /* extern "C" */
#endif

#endif  /* GSEVT_IO_H_ */
