// The KDTree3f wire and bag format: CDR serialization of
// kd_slam2_msgs/msg/KDTree3f and MCAP files (include/kdh_io.h).
//
// CDR (OMG CDR as ROS 2 serializes it, little endian): a 4-byte
// encapsulation header {0x00, 0x01, options}, then the fields in declaration
// order, each primitive aligned to its own size relative to the end of the
// header; strings as a uint32 length that counts the terminating NUL, the
// bytes and the NUL; fixed arrays as their elements; sequences as a uint32
// count and the elements.  The field order follows KDTree3f.msg and
// KDNode3f.msg (kd_slam2_msgs/msg/), with std_msgs/Header = {int32 sec,
// uint32 nanosec, string frame_id}.
//
// MCAP (https://mcap.dev/spec): magic, then records {uint8 opcode, uint64
// length, content}; strings as uint32 length + bytes; the reader walks the
// data section and uncompressed chunks, the writer emits Header, Schema,
// Channel and Message records, DataEnd, a summary section and the Footer.
#include <fcntl.h>
#include <stdio.h>
#include <string.h>
#include <sys/mman.h>
#include <sys/stat.h>
#include <unistd.h>

#include <map>
#include <string>
#include <vector>

#include "../../include/kdh.h"
#include "../../include/kdh_io.h"

namespace {

static_assert(sizeof(float) == 4 && sizeof(double) == 8, "IEEE sizes");

// ---- CDR ---------------------------------------------------------------------

// Writes (buf != nullptr) or only counts (buf == nullptr).
struct CdrOut {
  uint8_t* buf;
  size_t cap;
  size_t pos = 0;
  bool ok = true;
  CdrOut(uint8_t* b, size_t c) : buf(b), cap(c) {}
  void bytes(const void* p, size_t n) {
    if (buf) {
      if (pos + n > cap) {
        ok = false;
        return;
      }
      if (p)
        memcpy(buf + pos, p, n);
      else
        memset(buf + pos, 0, n);
    }
    pos += n;
  }
  void align(size_t a) {
    const size_t rel = pos - 4;  // alignment counts from the end of the encapsulation header
    const size_t pad = (a - rel % a) % a;
    if (pad) bytes(nullptr, pad);
  }
  template <typename T>
  void put(T v) {
    align(sizeof(T));
    bytes(&v, sizeof(T));
  }
  void str(const char* s) {
    const uint32_t n = (uint32_t)strlen(s) + 1;
    put<uint32_t>(n);
    bytes(s, n);
  }
};

struct CdrIn {
  const uint8_t* b;
  size_t len;
  size_t pos = 4;
  bool ok = true;
  CdrIn(const uint8_t* p, size_t n) : b(p), len(n) {}
  bool align(size_t a) {
    const size_t rel = pos - 4;
    const size_t pad = (a - rel % a) % a;
    if (pos + pad > len) return ok = false;
    pos += pad;
    return true;
  }
  template <typename T>
  bool get(T& v) {
    if (!align(sizeof(T)) || pos + sizeof(T) > len) return ok = false;
    memcpy(&v, b + pos, sizeof(T));
    pos += sizeof(T);
    return true;
  }
  bool str(const char** p, size_t* n) {
    uint32_t m = 0;
    if (!get(m) || m == 0 || pos + m > len) return ok = false;
    *p = (const char*)b + pos;
    *n = b[pos + m - 1] == 0 ? m - 1 : m;
    pos += m;
    return true;
  }
};

void encode_into(CdrOut& o, const kdh_kdtree3f_info& info, const kdh_node3f* nodes, int64_t n) {
  const uint8_t encapsulation[4] = {0x00, 0x01, 0x00, 0x00};  // CDR little endian, no options
  o.bytes(encapsulation, 4);
  o.put<int32_t>(info.stamp_sec);
  o.put<uint32_t>(info.stamp_nanosec);
  o.str(info.frame_id);
  o.str(info.src_topic);
  o.put<double>(info.ts_start);
  for (int i = 0; i < 3; ++i) o.put<float>(info.root_mean[i]);
  for (int i = 0; i < 3; ++i) o.put<float>(info.root_eigenvalues[i]);
  for (int i = 0; i < 9; ++i) o.put<float>(info.root_eigenvectors[i]);
  o.put<uint32_t>((uint32_t)n);
  for (int64_t i = 0; i < n; ++i) {
    const kdh_node3f& k = nodes ? nodes[i] : kdh_node3f{};
    for (int d = 0; d < 3; ++d) o.put<float>(k.mean[d]);
    for (int d = 0; d < 3; ++d) o.put<float>(k.direction[d]);
    o.put<uint32_t>((uint32_t)k.idx_start);  // static_cast<uint32_t> as the reference writer
    o.put<uint32_t>((uint32_t)k.idx_end);
    o.put<int32_t>(k.idx_left);
    o.put<int32_t>(k.idx_right);
    o.put<double>(k.dts_mean);
    o.put<double>(k.dts_covariance);
  }
}

// ---- MCAP --------------------------------------------------------------------

const uint8_t kMagic[8] = {0x89, 'M', 'C', 'A', 'P', '0', '\r', '\n'};

enum : uint8_t {
  kOpHeader = 0x01,
  kOpFooter = 0x02,
  kOpSchema = 0x03,
  kOpChannel = 0x04,
  kOpMessage = 0x05,
  kOpChunk = 0x06,
  kOpStatistics = 0x0B,
  kOpSummaryOffset = 0x0E,
  kOpDataEnd = 0x0F,
};

struct Span {
  const uint8_t* p;
  uint64_t n;
};

struct Rd {
  const uint8_t* p;
  uint64_t n, pos = 0;
  bool ok = true;
  template <typename T>
  T get() {
    T v{};
    if (pos + sizeof(T) > n) {
      ok = false;
      return v;
    }
    memcpy(&v, p + pos, sizeof(T));
    pos += sizeof(T);
    return v;
  }
  std::string str() {
    const uint32_t m = get<uint32_t>();
    if (!ok || pos + m > n) {
      ok = false;
      return std::string();
    }
    std::string s((const char*)p + pos, m);
    pos += m;
    return s;
  }
};

}  // namespace

struct kdh_mcap_reader {
  int fd = -1;
  const uint8_t* map = nullptr;
  size_t size = 0;
  uint64_t pos = 8;            // next record in the file
  Span chunk{nullptr, 0};      // records of the chunk being read
  uint64_t chunk_pos = 0;
  bool done = false;
  std::map<uint16_t, std::string> schema_names;
  struct Channel {
    std::string topic, schema, encoding;
  };
  std::map<uint16_t, Channel> channels;
  std::string err;
};

struct kdh_mcap_writer {
  FILE* f = nullptr;
  uint64_t pos = 0;
  struct Schema {
    uint16_t id;
    std::string name, text;
  };
  struct Channel {
    uint16_t id, schema_id;
    std::string topic;
    uint32_t seq = 0;
    uint64_t count = 0;
  };
  std::vector<Schema> schemas;
  std::vector<Channel> channels;
  uint64_t n_messages = 0, t_min = ~0ull, t_max = 0;
  bool ok = true;
};

namespace {

// One record {op, content} of a reader, from the chunk being read or the file.
bool next_record(kdh_mcap_reader* r, uint8_t* op, Span* content, bool* in_chunk) {
  if (r->chunk.p) {
    if (r->chunk_pos + 9 <= r->chunk.n) {
      const uint8_t* q = r->chunk.p + r->chunk_pos;
      uint64_t len;
      memcpy(&len, q + 1, 8);
      if (len > r->chunk.n - r->chunk_pos - 9) {
        r->err = "truncated record inside a chunk";
        return false;
      }
      *op = q[0];
      *content = Span{q + 9, len};
      r->chunk_pos += 9 + len;
      *in_chunk = true;
      return true;
    }
    r->chunk = Span{nullptr, 0};
  }
  if (r->pos + 9 > r->size) {
    r->err = "file ends without a footer";
    return false;
  }
  const uint8_t* q = r->map + r->pos;
  uint64_t len;
  memcpy(&len, q + 1, 8);
  if (len > r->size - r->pos - 9) {
    r->err = "truncated record";
    return false;
  }
  *op = q[0];
  *content = Span{q + 9, len};
  r->pos += 9 + len;
  *in_chunk = false;
  return true;
}

void put_bytes(kdh_mcap_writer* w, const void* p, size_t n) {
  if (n && fwrite(p, 1, n, w->f) != n) w->ok = false;
  w->pos += n;
}

struct Rec {
  std::vector<uint8_t> b;
  template <typename T>
  void put(T v) {
    const uint8_t* p = (const uint8_t*)&v;
    b.insert(b.end(), p, p + sizeof(T));
  }
  void str(const std::string& s) {
    put<uint32_t>((uint32_t)s.size());
    b.insert(b.end(), s.begin(), s.end());
  }
  void raw(const void* p, size_t n) { b.insert(b.end(), (const uint8_t*)p, (const uint8_t*)p + n); }
};

void put_record(kdh_mcap_writer* w, uint8_t op, const Rec& r) {
  put_bytes(w, &op, 1);
  const uint64_t len = r.b.size();
  put_bytes(w, &len, 8);
  put_bytes(w, r.b.data(), r.b.size());
}

Rec schema_record(const kdh_mcap_writer::Schema& s) {
  Rec r;
  r.put<uint16_t>(s.id);
  r.str(s.name);
  r.str("ros2msg");
  r.put<uint32_t>((uint32_t)s.text.size());
  r.raw(s.text.data(), s.text.size());
  return r;
}

Rec channel_record(const kdh_mcap_writer::Channel& c) {
  Rec r;
  r.put<uint16_t>(c.id);
  r.put<uint16_t>(c.schema_id);
  r.str(c.topic);
  r.str("cdr");
  r.put<uint32_t>(0);  // empty metadata map
  return r;
}

const char kSchemaKDTree3f[] =
    "std_msgs/Header header\n"
    "string src_topic\n"
    "float64 ts_start\n"
    "float32[3] root_mean\n"
    "float32[3] root_eigenvalues\n"
    "float32[9] root_eigenvectors\n"
    "kd_slam2_msgs/KDNode3f[] nodes\n"
    "================================================================================\n"
    "MSG: std_msgs/Header\n"
    "builtin_interfaces/Time stamp\n"
    "string frame_id\n"
    "================================================================================\n"
    "MSG: builtin_interfaces/Time\n"
    "int32 sec\n"
    "uint32 nanosec\n"
    "================================================================================\n"
    "MSG: kd_slam2_msgs/KDNode3f\n"
    "float32[3] mean\n"
    "float32[3] direction\n"
    "uint32 idx_start\n"
    "uint32 idx_end\n"
    "int32 idx_left\n"
    "int32 idx_right\n"
    "float64 dts_mean\n"
    "float64 dts_covariance\n";

}  // namespace

extern "C" {

// ---- KDTree3f ------------------------------------------------------------------

int kdh_kdtree3f_encoded_size(const kdh_kdtree3f_info* info, int64_t n_nodes, size_t* size) {
  if (!info || !size || !info->frame_id || !info->src_topic || n_nodes < 0 || n_nodes > (int64_t)0xffffffff)
    return KDH_ERR_INVALID;
  CdrOut o(nullptr, 0);
  encode_into(o, *info, nullptr, n_nodes);
  *size = o.pos;
  return KDH_OK;
}

int kdh_kdtree3f_encode(const kdh_kdtree3f_info* info, const kdh_node3f* nodes, int64_t n_nodes, uint8_t* buf,
                        size_t cap, size_t* len) {
  if (!info || !buf || !len || !info->frame_id || !info->src_topic || n_nodes < 0 ||
      n_nodes > (int64_t)0xffffffff || (n_nodes > 0 && !nodes))
    return KDH_ERR_INVALID;
  CdrOut o(buf, cap);
  encode_into(o, *info, nodes, n_nodes);
  if (!o.ok) return KDH_ERR_INVALID;
  *len = o.pos;
  return KDH_OK;
}

int kdh_kdtree3f_decode(const uint8_t* buf, size_t len, kdh_kdtree3f_info* info, char* frame_id, size_t frame_cap,
                        char* src_topic, size_t topic_cap, kdh_node3f* nodes, int64_t node_cap, int64_t* n_nodes) {
  if (!buf || !info || !n_nodes || len < 4) return KDH_ERR_INVALID;
  if (buf[0] != 0x00 || buf[1] != 0x01) return KDH_ERR_INVALID;  // CDR little endian only
  CdrIn in(buf, len);
  kdh_kdtree3f_info out;
  memset(&out, 0, sizeof(out));
  const char* s;
  size_t sn;
  in.get(out.stamp_sec);
  in.get(out.stamp_nanosec);
  auto copy_str = [&](char* dst, size_t cap) {
    if (!in.str(&s, &sn)) return (const char*)nullptr;
    if (dst && cap) {
      const size_t m = sn < cap - 1 ? sn : cap - 1;
      memcpy(dst, s, m);
      dst[m] = 0;
    }
    return (const char*)dst;
  };
  out.frame_id = copy_str(frame_id, frame_cap);
  out.src_topic = copy_str(src_topic, topic_cap);
  in.get(out.ts_start);
  for (int i = 0; i < 3; ++i) in.get(out.root_mean[i]);
  for (int i = 0; i < 3; ++i) in.get(out.root_eigenvalues[i]);
  for (int i = 0; i < 9; ++i) in.get(out.root_eigenvectors[i]);
  uint32_t n = 0;
  in.get(n);
  if (!in.ok) return KDH_ERR_INVALID;
  // every node takes at least 52 bytes
  if ((uint64_t)n * 52 > (uint64_t)(len - in.pos) + 4) return KDH_ERR_INVALID;
  *n_nodes = n;
  const bool fill = nodes && node_cap >= (int64_t)n;
  // A node is 40 bytes of 4-byte fields (mean, direction, idx_start, idx_end,
  // idx_left, idx_right) and two doubles at 8-byte alignment: the first may
  // carry 4 bytes of padding, every later one starts 8-aligned and takes 56.
  // The end of the records is therefore known before reading them, and a
  // shorter buffer is the only way they can be malformed.
  const size_t first = in.pos;
  const size_t pad = (8 - (first + 40 - 4) % 8) % 8;
  const size_t end = n > 0 ? first + 40 + pad + 16 + (size_t)56 * (n - 1) : first;
  if (end > len) return KDH_ERR_INVALID;
  if (fill) {
    const uint8_t* p = buf + first;
    for (uint32_t i = 0; i < n; ++i) {
      kdh_node3f k;
      memset(&k, 0, sizeof(k));
      uint32_t a = 0, b = 0;
      memcpy(k.mean, p, 12);
      memcpy(k.direction, p + 12, 12);
      memcpy(&a, p + 24, 4);
      memcpy(&b, p + 28, 4);
      memcpy(&k.idx_left, p + 32, 4);
      memcpy(&k.idx_right, p + 36, 4);
      const uint8_t* q = p + 40 + (i == 0 ? pad : 0);
      memcpy(&k.dts_mean, q, 8);
      memcpy(&k.dts_covariance, q + 8, 8);
      k.idx_start = a;
      k.idx_end = b;
      nodes[i] = k;
      p = q + 16;
    }
  }
  *info = out;
  return KDH_OK;
}

int kdh_tree_leaves(const kdh_node3f* nodes, int64_t n_nodes, uint64_t* leaves, int64_t cap, int64_t* n_leaves) {
  if (!n_leaves || n_nodes < 0 || (n_nodes > 0 && !nodes)) return KDH_ERR_INVALID;
  int64_t k = 0;
  for (int64_t i = 0; i < n_nodes; ++i)
    if (nodes[i].idx_left < 0 && nodes[i].idx_right < 0) {
      if (leaves && k < cap) leaves[k] = (uint64_t)i;
      ++k;
    }
  *n_leaves = k;
  return KDH_OK;
}

const char* kdh_kdtree3f_schema_text(void) { return kSchemaKDTree3f; }

// ---- MCAP reader ---------------------------------------------------------------

int kdh_mcap_reader_open(const char* path, kdh_mcap_reader** out) {
  if (!path || !out) return KDH_ERR_INVALID;
  *out = nullptr;
  const int fd = open(path, O_RDONLY);
  if (fd < 0) return KDH_ERR_NOT_FOUND;
  struct stat st;
  if (fstat(fd, &st) != 0 || st.st_size < 16) {
    close(fd);
    return KDH_ERR_INVALID;
  }
  void* m = mmap(nullptr, (size_t)st.st_size, PROT_READ, MAP_PRIVATE, fd, 0);
  if (m == MAP_FAILED) {
    close(fd);
    return KDH_ERR_INVALID;
  }
  if (memcmp(m, kMagic, 8) != 0) {
    munmap(m, (size_t)st.st_size);
    close(fd);
    return KDH_ERR_INVALID;
  }
  auto* r = new kdh_mcap_reader();
  r->fd = fd;
  r->map = (const uint8_t*)m;
  r->size = (size_t)st.st_size;
  *out = r;
  return KDH_OK;
}

int kdh_mcap_reader_next(kdh_mcap_reader* r, kdh_mcap_message* msg) {
  if (!r || !msg) return KDH_ERR_INVALID;
  while (!r->done) {
    uint8_t op;
    Span c;
    bool in_chunk;
    if (!next_record(r, &op, &c, &in_chunk)) return KDH_ERR_INVALID;
    Rd d{c.p, c.n};
    switch (op) {
      case kOpSchema: {
        const uint16_t id = d.get<uint16_t>();
        std::string name = d.str();
        if (!d.ok) {
          r->err = "malformed schema record";
          return KDH_ERR_INVALID;
        }
        r->schema_names[id] = name;
        break;
      }
      case kOpChannel: {
        const uint16_t id = d.get<uint16_t>();
        const uint16_t sid = d.get<uint16_t>();
        std::string topic = d.str();
        std::string enc = d.str();
        if (!d.ok) {
          r->err = "malformed channel record";
          return KDH_ERR_INVALID;
        }
        kdh_mcap_reader::Channel& ch = r->channels[id];
        ch.topic = topic;
        ch.encoding = enc;
        ch.schema = sid && r->schema_names.count(sid) ? r->schema_names[sid] : std::string();
        break;
      }
      case kOpMessage: {
        const uint16_t cid = d.get<uint16_t>();
        const uint32_t seq = d.get<uint32_t>();
        const uint64_t lt = d.get<uint64_t>();
        const uint64_t pt = d.get<uint64_t>();
        if (!d.ok || !r->channels.count(cid)) {
          r->err = "message record without a known channel";
          return KDH_ERR_INVALID;
        }
        const kdh_mcap_reader::Channel& ch = r->channels[cid];
        msg->topic = ch.topic.c_str();
        msg->schema_name = ch.schema.c_str();
        msg->message_encoding = ch.encoding.c_str();
        msg->log_time = lt;
        msg->publish_time = pt;
        msg->sequence = seq;
        msg->channel_id = cid;
        msg->data = c.p + d.pos;
        msg->size = c.n - d.pos;
        return KDH_OK;
      }
      case kOpChunk: {
        if (in_chunk) {
          r->err = "chunk inside a chunk";
          return KDH_ERR_INVALID;
        }
        d.get<uint64_t>();  // message start time
        d.get<uint64_t>();  // message end time
        const uint64_t usize = d.get<uint64_t>();
        d.get<uint32_t>();  // uncompressed crc
        std::string comp = d.str();
        const uint64_t rlen = d.get<uint64_t>();
        if (!d.ok || rlen > c.n - d.pos) {
          r->err = "malformed chunk record";
          return KDH_ERR_INVALID;
        }
        if (!comp.empty()) {
          r->err = "compressed chunk (" + comp + ") is not supported";
          return KDH_ERR_INVALID;
        }
        if (usize != rlen) {
          r->err = "uncompressed chunk size mismatch";
          return KDH_ERR_INVALID;
        }
        r->chunk = Span{c.p + d.pos, rlen};
        r->chunk_pos = 0;
        break;
      }
      case kOpFooter:
        r->done = true;
        break;
      default:
        break;  // indexes, statistics, attachments, metadata, data end
    }
  }
  return KDH_ERR_NOT_FOUND;
}

const char* kdh_mcap_reader_error(const kdh_mcap_reader* r) { return r ? r->err.c_str() : "null reader"; }

int kdh_mcap_reader_close(kdh_mcap_reader* r) {
  if (!r) return KDH_ERR_INVALID;
  munmap((void*)r->map, r->size);
  close(r->fd);
  delete r;
  return KDH_OK;
}

// ---- MCAP writer -----------------------------------------------------------------

int kdh_mcap_writer_open(const char* path, kdh_mcap_writer** out) {
  if (!path || !out) return KDH_ERR_INVALID;
  *out = nullptr;
  FILE* f = fopen(path, "wb");
  if (!f) return KDH_ERR_INVALID;
  auto* w = new kdh_mcap_writer();
  w->f = f;
  put_bytes(w, kMagic, 8);
  Rec h;
  h.str("ros2");
  h.str("kd_slam2_amd");
  put_record(w, kOpHeader, h);
  if (!w->ok) {
    fclose(f);
    delete w;
    return KDH_ERR_INVALID;
  }
  *out = w;
  return KDH_OK;
}

int kdh_mcap_writer_add_channel(kdh_mcap_writer* w, const char* topic, const char* schema_name,
                                const char* schema_text, uint16_t* channel_id) {
  if (!w || !topic || !schema_name || !schema_text || !channel_id) return KDH_ERR_INVALID;
  uint16_t sid = 0;
  for (const auto& s : w->schemas)
    if (s.name == schema_name && s.text == schema_text) sid = s.id;
  if (!sid) {
    kdh_mcap_writer::Schema s{(uint16_t)(w->schemas.size() + 1), schema_name, schema_text};
    w->schemas.push_back(s);
    put_record(w, kOpSchema, schema_record(s));
    sid = s.id;
  }
  kdh_mcap_writer::Channel c;
  c.id = (uint16_t)w->channels.size();
  c.schema_id = sid;
  c.topic = topic;
  w->channels.push_back(c);
  put_record(w, kOpChannel, channel_record(c));
  *channel_id = c.id;
  return w->ok ? KDH_OK : KDH_ERR_INVALID;
}

int kdh_mcap_writer_write(kdh_mcap_writer* w, uint16_t channel_id, uint64_t log_time, uint64_t publish_time,
                          const uint8_t* data, uint64_t size) {
  if (!w || (size && !data) || channel_id >= w->channels.size()) return KDH_ERR_INVALID;
  kdh_mcap_writer::Channel& c = w->channels[channel_id];
  Rec m;
  m.put<uint16_t>(channel_id);
  m.put<uint32_t>(c.seq++);
  m.put<uint64_t>(log_time);
  m.put<uint64_t>(publish_time);
  const uint8_t op = kOpMessage;
  const uint64_t len = m.b.size() + size;
  put_bytes(w, &op, 1);
  put_bytes(w, &len, 8);
  put_bytes(w, m.b.data(), m.b.size());
  put_bytes(w, data, size);
  c.count++;
  w->n_messages++;
  if (log_time < w->t_min) w->t_min = log_time;
  if (log_time > w->t_max) w->t_max = log_time;
  return w->ok ? KDH_OK : KDH_ERR_INVALID;
}

int kdh_mcap_writer_close(kdh_mcap_writer* w) {
  if (!w) return KDH_ERR_INVALID;
  Rec de;
  de.put<uint32_t>(0);  // data section crc: not computed
  put_record(w, kOpDataEnd, de);
  const uint64_t summary_start = w->pos;
  struct Group {
    uint8_t op;
    uint64_t start, length;
  };
  std::vector<Group> groups;
  if (!w->schemas.empty()) {
    const uint64_t s0 = w->pos;
    for (const auto& s : w->schemas) put_record(w, kOpSchema, schema_record(s));
    groups.push_back({kOpSchema, s0, w->pos - s0});
  }
  if (!w->channels.empty()) {
    const uint64_t s0 = w->pos;
    for (const auto& c : w->channels) put_record(w, kOpChannel, channel_record(c));
    groups.push_back({kOpChannel, s0, w->pos - s0});
  }
  {
    const uint64_t s0 = w->pos;
    Rec st;
    st.put<uint64_t>(w->n_messages);
    st.put<uint16_t>((uint16_t)w->schemas.size());
    st.put<uint32_t>((uint32_t)w->channels.size());
    st.put<uint32_t>(0);  // attachments
    st.put<uint32_t>(0);  // metadata
    st.put<uint32_t>(0);  // chunks
    st.put<uint64_t>(w->n_messages ? w->t_min : 0);
    st.put<uint64_t>(w->n_messages ? w->t_max : 0);
    Rec counts;
    for (const auto& c : w->channels) {
      counts.put<uint16_t>(c.id);
      counts.put<uint64_t>(c.count);
    }
    st.put<uint32_t>((uint32_t)counts.b.size());
    st.raw(counts.b.data(), counts.b.size());
    put_record(w, kOpStatistics, st);
    groups.push_back({kOpStatistics, s0, w->pos - s0});
  }
  const uint64_t offsets_start = w->pos;
  for (const Group& g : groups) {
    Rec so;
    so.put<uint8_t>(g.op);
    so.put<uint64_t>(g.start);
    so.put<uint64_t>(g.length);
    put_record(w, kOpSummaryOffset, so);
  }
  Rec ft;
  ft.put<uint64_t>(summary_start);
  ft.put<uint64_t>(offsets_start);
  ft.put<uint32_t>(0);  // summary crc: not computed
  put_record(w, kOpFooter, ft);
  put_bytes(w, kMagic, 8);
  const bool ok = w->ok && fclose(w->f) == 0;
  delete w;
  return ok ? KDH_OK : KDH_ERR_INVALID;
}

}  // extern "C"
