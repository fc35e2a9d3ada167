// Snapshots of a lattice (C-ABI lbm_snapshot_save / lbm_snapshot_load):
// the whole state needed to resume, in one binary file.
//
//   header   magic "LBMSNAP1", version, lbm_config, steps done, section sizes
//   f        the current population buffer as stored on the device (blocked
//            layout, storage type): restores bit-exactly
//   standard mode: link mask (uint32 per node), class map (2 bits per node)
//   reference mode: flag bytes, reflection records (nodes, 14 direction
//            words and 81 weights each), flow injection (kind, parameters,
//            induced vector, mask bytes when the flow selects by mask), the
//            force register
//
// The reference keeps its state in a device array of 144-byte records that
// retrieve_data / load_data (Lattice.cpp:73-113) copy to and from the
// host; a snapshot holds the same information for either mode without
// the per-node record overhead.
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstring>
#include <vector>

#include "lbm_internal.h"
#include "lbm_layout.h"
#include "lbm_store.h"

using namespace lbm;

namespace {

constexpr char kMagic[8] = {'L', 'B', 'M', 'S', 'N', 'A', 'P', '1'};
constexpr uint32_t kVersion = 2;  // 2: FP16 times 2^12, standard FP32 shifted

struct SnapHeader {
  char magic[8];
  uint32_t version;
  uint32_t header_bytes;
  lbm_config cfg;
  uint64_t steps_done;
  uint64_t pop_bytes;
  uint64_t mask_bytes;
  uint64_t cls_bytes;
  uint64_t flag_bytes;
  int32_t nrec;
  int32_t flow_kind;
  int32_t flow_params[6];
  float flow_induced[27];
  uint64_t flow_mask_bytes;
  float force[3];
  uint32_t reserved;
};

struct File {
  FILE* f = nullptr;
  ~File() {
    if (f) std::fclose(f);
  }
};

int write_all(FILE* f, const void* p, size_t n) {
  if (n && std::fwrite(p, 1, n, f) != n) return fail(LBM_ERR_INVALID, "snapshot write failed");
  return LBM_OK;
}

int read_all(FILE* f, void* p, size_t n) {
  if (n && std::fread(p, 1, n, f) != n)
    return fail(LBM_ERR_INVALID, "snapshot file is truncated");
  return LBM_OK;
}

int download(lbm_ctx* c, void* dst, const void* src, size_t n) {
  if (!n) return LBM_OK;
  LBM_HIP(hipMemcpyAsync(dst, src, n, hipMemcpyDeviceToHost, c->stream));
  LBM_HIP(hipStreamSynchronize(c->stream));
  return LBM_OK;
}

int upload(lbm_ctx* c, void* dst, const void* src, size_t n) {
  if (!n) return LBM_OK;
  LBM_HIP(hipMemcpyAsync(dst, src, n, hipMemcpyHostToDevice, c->stream));
  LBM_HIP(hipStreamSynchronize(c->stream));
  return LBM_OK;
}

// Bytes of the file after the header, or -1 when its size is unknown.
int64_t bytes_after_header(FILE* f) {
  const long here = std::ftell(f);
  if (here < 0 || std::fseek(f, 0, SEEK_END) != 0) return -1;
  const long end = std::ftell(f);
  if (end < 0 || std::fseek(f, here, SEEK_SET) != 0) return -1;
  return (int64_t)end - (int64_t)here;
}

// The checks that lbm_upload_reflections, lbm_set_flow and the receiver
// tables of the fused reference step apply to the same data, made before
// the context changes: every record's node in range and recorded once, a
// reflecting node with a record inside the obstacle, a known flow kind.
int check_reference_sections(uint64_t N, const std::vector<uint8_t>& flags,
                             const std::vector<int32_t>& rnode, const SnapHeader& h) {
  std::vector<uint8_t> seen(N, 0);
  for (const int32_t node : rnode) {
    if (node < 0 || (uint64_t)node >= N)
      return fail(LBM_ERR_INVALID, "snapshot reflection record node out of range");
    if (seen[node]) return fail(LBM_ERR_INVALID, "snapshot holds two reflection records for a node");
    seen[node] = 1;
    if ((flags[node] & LBM_FLAG_REFLECTED) && !(flags[node] & LBM_FLAG_INTERNAL))
      return fail(LBM_ERR_INVALID, "snapshot reflecting node is not internal");
  }
  if (h.flow_kind < LBM_FLOW_NONE || h.flow_kind > LBM_FLOW_MASK)
    return fail(LBM_ERR_INVALID, "snapshot flow kind unknown");
  if (h.flow_kind == LBM_FLOW_PLANE && (h.flow_params[0] < 0 || h.flow_params[0] > 2))
    return fail(LBM_ERR_INVALID, "snapshot flow plane axis must be 0..2");
  return LBM_OK;
}

}  // namespace

// Entry points of include/lbm.h (exported by lbm_exports.hip).
namespace lbm {
namespace api {

int lbm_snapshot_save(lbm_ctx* c, const char* path) {
  if (!c || !path) return fail(LBM_ERR_INVALID, "null argument");
  const uint64_t N = c->grid.n;
  const bool standard = c->cfg.mode == LBM_MODE_STANDARD;
  SnapHeader h{};
  std::memcpy(h.magic, kMagic, sizeof(kMagic));
  h.version = kVersion;
  h.header_bytes = sizeof(SnapHeader);
  h.cfg = c->cfg;
  h.steps_done = c->steps_done;
  h.pop_bytes = population_elems(c->q, N) * storage_bytes(c->cfg.storage);
  h.mask_bytes = standard ? 4 * N : 0;
  h.cls_bytes = standard ? 4 * ((N + 15) / 16) : 0;
  h.flag_bytes = standard ? 0 : N;
  h.nrec = standard ? 0 : c->nrec;
  h.flow_kind = c->flow.kind;
  std::memcpy(h.flow_params, c->flow.params, sizeof(h.flow_params));
  std::memcpy(h.flow_induced, c->flow.induced, sizeof(h.flow_induced));
  h.flow_mask_bytes = (!standard && c->flow.kind == LBM_FLOW_MASK) ? N : 0;
  int rc;
  std::vector<uint8_t> buf(h.pop_bytes);
  if ((rc = download(c, buf.data(), c->f[c->cur], h.pop_bytes))) return rc;
  if (!standard && (rc = download(c, h.force, c->force, sizeof(h.force)))) return rc;
  File out;
  out.f = std::fopen(path, "wb");
  if (!out.f) return fail(LBM_ERR_INVALID, std::string("cannot open ") + path);
  if ((rc = write_all(out.f, &h, sizeof(h))) || (rc = write_all(out.f, buf.data(), buf.size())))
    return rc;
  if (standard) {
    buf.resize(h.mask_bytes);
    if ((rc = download(c, buf.data(), c->mask, h.mask_bytes)) ||
        (rc = write_all(out.f, buf.data(), h.mask_bytes)))
      return rc;
    buf.resize(h.cls_bytes);
    if ((rc = download(c, buf.data(), c->cls, h.cls_bytes)) ||
        (rc = write_all(out.f, buf.data(), h.cls_bytes)))
      return rc;
  } else {
    if ((rc = write_all(out.f, c->host_flags.data(), N))) return rc;
    if (h.nrec > 0 &&
        ((rc = write_all(out.f, c->host_rec_node.data(), 4 * (size_t)h.nrec)) ||
         (rc = write_all(out.f, c->host_rec_dirs.data(), 4 * 14 * (size_t)h.nrec)) ||
         (rc = write_all(out.f, c->host_rec_w.data(), 4 * 81 * (size_t)h.nrec))))
      return rc;
    if (h.flow_mask_bytes) {
      buf.resize(N);
      if ((rc = download(c, buf.data(), c->flow_mask, N)) || (rc = write_all(out.f, buf.data(), N)))
        return rc;
    }
  }
  if (std::fclose(out.f) != 0) {
    out.f = nullptr;
    return fail(LBM_ERR_INVALID, "snapshot write failed");
  }
  out.f = nullptr;
  return LBM_OK;
}

int lbm_snapshot_load(lbm_ctx* c, const char* path) {
  if (!c || !path) return fail(LBM_ERR_INVALID, "null argument");
  File in;
  in.f = std::fopen(path, "rb");
  if (!in.f) return fail(LBM_ERR_INVALID, std::string("cannot open ") + path);
  SnapHeader h{};
  int rc;
  if ((rc = read_all(in.f, &h, sizeof(h)))) return rc;
  if (std::memcmp(h.magic, kMagic, sizeof(kMagic)) != 0 || h.version != kVersion ||
      h.header_bytes != sizeof(SnapHeader))
    return fail(LBM_ERR_INVALID, "not a snapshot of this library version");
  const lbm_config& s = h.cfg;
  if (s.nx != c->cfg.nx || s.ny != c->cfg.ny || s.nz != c->cfg.nz || s.lattice != c->cfg.lattice ||
      s.storage != c->cfg.storage || s.mode != c->cfg.mode)
    return fail(LBM_ERR_INVALID,
                "snapshot extents / lattice / storage / mode differ from the context");
  // the snapshot's parameters replace the context's and pass the checks a
  // new context's would (the geometry tables come with the snapshot)
  if (int e = check_config(&s)) return e;
  const uint64_t N = c->grid.n;
  const bool standard = c->cfg.mode == LBM_MODE_STANDARD;
  // a node holds at most one reflection record, so nrec <= N; standard
  // lattices have none
  if (h.pop_bytes != population_elems(c->q, N) * storage_bytes(c->cfg.storage) ||
      h.mask_bytes != (standard ? 4 * N : 0) || h.cls_bytes != (standard ? 4 * ((N + 15) / 16) : 0) ||
      h.flag_bytes != (standard ? 0 : N) || h.nrec < 0 || (uint64_t)h.nrec > N ||
      (standard && h.nrec != 0) ||
      h.flow_mask_bytes != ((!standard && h.flow_kind == LBM_FLOW_MASK) ? N : 0))
    return fail(LBM_ERR_INVALID, "snapshot sections do not match the context");
  // the sections must fill the rest of the file exactly: checked before
  // any buffer is sized from the header
  const uint64_t rec_bytes = (uint64_t)h.nrec * (4 + 4 * 14 + 4 * 81);
  const uint64_t body = h.pop_bytes + h.mask_bytes + h.cls_bytes + h.flag_bytes + rec_bytes +
                        h.flow_mask_bytes;
  const int64_t left = bytes_after_header(in.f);
  if (left < 0 || (uint64_t)left != body)
    return fail(LBM_ERR_INVALID, "snapshot file size does not match its header");
  std::vector<uint8_t> pops(h.pop_bytes), mask, cls, flags, fmask;
  std::vector<int32_t> rnode;
  std::vector<uint32_t> rdirs;
  std::vector<float> rw;
  if ((rc = read_all(in.f, pops.data(), pops.size()))) return rc;
  if (standard) {
    mask.resize(h.mask_bytes);
    cls.resize(h.cls_bytes);
    if ((rc = read_all(in.f, mask.data(), mask.size())) || (rc = read_all(in.f, cls.data(), cls.size())))
      return rc;
  } else {
    flags.resize(N);
    rnode.resize(h.nrec);
    rdirs.resize(14 * (size_t)h.nrec);
    rw.resize(81 * (size_t)h.nrec);
    if ((rc = read_all(in.f, flags.data(), N)) ||
        (rc = read_all(in.f, rnode.data(), 4 * rnode.size())) ||
        (rc = read_all(in.f, rdirs.data(), 4 * rdirs.size())) ||
        (rc = read_all(in.f, rw.data(), 4 * rw.size())))
      return rc;
    if (h.flow_mask_bytes) {
      fmask.resize(N);
      if ((rc = read_all(in.f, fmask.data(), N))) return rc;
    }
    if ((rc = check_reference_sections(N, flags, rnode, h))) return rc;
  }
  // everything read and checked: now change the context
  c->cfg.tau = s.tau;
  for (int a = 0; a < 3; ++a) {
    c->cfg.force[a] = s.force[a];
    c->cfg.lid_u[a] = s.lid_u[a];
    c->cfg.periodic[a] = s.periodic[a];
  }
  c->cfg.lid_face = s.lid_face;
  c->cfg.wall_rho = s.wall_rho;
  bool any_solid = false;
  for (size_t n = 0; standard && n < N && !any_solid; ++n) {
    uint32_t m;
    std::memcpy(&m, mask.data() + 4 * n, 4);
    any_solid = (m & LBM_MASK_SOLID) != 0u;
  }
  c->grid.flags = wrap_bits(c->cfg.periodic) | (any_solid ? kReadMasks : 0u);
  ++c->params_version;
  c->steps_done = h.steps_done;
  if ((rc = upload(c, c->f[c->cur], pops.data(), pops.size()))) return rc;
  LBM_HIP(hipMemcpyAsync(c->f[c->cur ^ 1], c->f[c->cur], pops.size(), hipMemcpyDeviceToDevice,
                         c->stream));
  if (standard) {
    if ((rc = upload(c, c->mask, mask.data(), mask.size())) ||
        (rc = upload(c, c->cls, cls.data(), cls.size())) ||
        (has_mask16(c->cfg) && (rc = launch_mask16(c->grid, c->mask, c->stream))))
      return rc;
  } else {
    if ((rc = lbm_upload_flags(c, flags.data()))) return rc;
    if ((rc = lbm_upload_reflections(c, h.nrec, rnode.data(), rdirs.data(), rw.data()))) return rc;
    if ((rc = lbm_set_flow(c, h.flow_kind, h.flow_params, h.flow_induced,
                           h.flow_mask_bytes ? fmask.data() : nullptr)))
      return rc;
    if ((rc = lbm_set_force(c, h.force))) return rc;
  }
  LBM_HIP(hipStreamSynchronize(c->stream));
  return LBM_OK;
}

}  // namespace api
}  // namespace lbm
