/*
 * kdh_io.h -- the KDTree3f wire and bag format (host code in libkdhip.so).
 *
 * A tree travels between processes as the ROS 2 message
 * kd_slam2_msgs/msg/KDTree3f (kd_slam2_msgs/msg/KDTree3f.msg, KDNode3f.msg),
 * serialized as little-endian CDR, and is stored in rosbag2 bags, whose
 * default storage on the reference's ROS distribution is MCAP.  The
 * reference converts with rclcpp::Serialization and rosbag2_cpp
 * (kd_slam2/src/lib/kd_io/rosbag2_message_writer.cpp:257-286,
 * rosbag2_message_reader.cpp:150-190); neither library is needed here.
 *
 * Node conversion (the reference's): mean and direction copied, idx_start and
 * idx_end narrowed to uint32 (static_cast), idx_left / idx_right int32,
 * dts_mean and dts_covariance float64.  The message carries no leaf list;
 * kdh_tree_leaves recomputes it (TreeCPU_::recomputeLeaves,
 * tree/tree_cpu_.h:194-205) for kdh_tree_upload.
 */
#ifndef KDH_IO_H_
#define KDH_IO_H_

#include <stddef.h>
#include <stdint.h>

#include "kdh.h"

#ifdef __cplusplus
extern "C" {
#endif

/* The non-node fields of KDTree3f.  root_eigenvectors is column-major, as in
 * the message (element col * 3 + row). */
typedef struct kdh_kdtree3f_info {
  int32_t stamp_sec;          /* header.stamp */
  uint32_t stamp_nanosec;
  const char* frame_id;       /* header.frame_id (encode: input; decode: the caller's buffer) */
  const char* src_topic;
  double ts_start;
  float root_mean[3];
  float root_eigenvalues[3];
  float root_eigenvectors[9];
} kdh_kdtree3f_info;

/* Serialized size (4-byte encapsulation header included) of a message with
 * these strings and n_nodes nodes. */
int kdh_kdtree3f_encoded_size(const kdh_kdtree3f_info* info, int64_t n_nodes, size_t* size);

/* Serializes info + nodes (Node_ records) into buf (cap bytes); *len gets
 * the size.  KDH_ERR_INVALID if cap is too small or a string is NULL. */
int kdh_kdtree3f_encode(const kdh_kdtree3f_info* info, const kdh_node3f* nodes, int64_t n_nodes, uint8_t* buf,
                        size_t cap, size_t* len);

/* Parses a serialized message.  *n_nodes always gets the node count; nodes
 * are written when node_cap suffices (pass NULL / 0 to ask for the count).
 * The strings are copied into frame_id / src_topic (NUL-terminated, cut to
 * the capacities) and info's pointers set to them.  KDH_ERR_INVALID for a
 * truncated or malformed buffer, or a representation other than CDR little
 * endian. */
int kdh_kdtree3f_decode(const uint8_t* buf, size_t len, kdh_kdtree3f_info* info, char* frame_id, size_t frame_cap,
                        char* src_topic, size_t topic_cap, kdh_node3f* nodes, int64_t node_cap, int64_t* n_nodes);

/* recomputeLeaves: indices of the nodes without children, in node order.
 * *n_leaves gets the count; leaves written when cap suffices. */
int kdh_tree_leaves(const kdh_node3f* nodes, int64_t n_nodes, uint64_t* leaves, int64_t cap, int64_t* n_leaves);

/* ---- MCAP files (https://mcap.dev/spec) ----------------------------------
 * Reader: every Message record in file order, from the data section and from
 * uncompressed Chunk records (a compressed chunk is an error, named in
 * kdh_mcap_reader_error).  The file is mapped; message data stays valid until
 * kdh_mcap_reader_close. */
typedef struct kdh_mcap_reader kdh_mcap_reader;

typedef struct kdh_mcap_message {
  const char* topic;
  const char* schema_name;       /* e.g. kd_slam2_msgs/msg/KDTree3f */
  const char* message_encoding;  /* e.g. cdr */
  uint64_t log_time;             /* ns */
  uint64_t publish_time;
  uint32_t sequence;
  uint16_t channel_id;
  const uint8_t* data;
  uint64_t size;
} kdh_mcap_message;

int kdh_mcap_reader_open(const char* path, kdh_mcap_reader** out);
/* KDH_OK with *msg filled, KDH_ERR_NOT_FOUND after the last message. */
int kdh_mcap_reader_next(kdh_mcap_reader* r, kdh_mcap_message* msg);
const char* kdh_mcap_reader_error(const kdh_mcap_reader* r);
int kdh_mcap_reader_close(kdh_mcap_reader* r);

/* Writer: profile "ros2", schemas "ros2msg", messages "cdr", Message records
 * in the data section, then a summary section (schemas, channels,
 * statistics, summary offsets) and the footer.  CRC fields are 0 ("not
 * computed").  No rosbag2 metadata.yaml is written (ros2 bag reindex makes
 * one). */
typedef struct kdh_mcap_writer kdh_mcap_writer;

int kdh_mcap_writer_open(const char* path, kdh_mcap_writer** out);
/* schema_text: the ros2msg definition (the type's fields, then each
 * dependency after a line of '=' and "MSG: pkg/Type"). */
int kdh_mcap_writer_add_channel(kdh_mcap_writer* w, const char* topic, const char* schema_name,
                                const char* schema_text, uint16_t* channel_id);
int kdh_mcap_writer_write(kdh_mcap_writer* w, uint16_t channel_id, uint64_t log_time, uint64_t publish_time,
                          const uint8_t* data, uint64_t size);
int kdh_mcap_writer_close(kdh_mcap_writer* w);

/* The ros2msg definition of kd_slam2_msgs/msg/KDTree3f with its dependencies
 * (a static string). */
const char* kdh_kdtree3f_schema_text(void);

#ifdef __cplusplus
}
#endif

#endif /* KDH_IO_H_ */
