// Restart files of SolverTest's LocalScalar3D
// (examples/SolverTest/LocalScalar3D.cpp:84-357).
//
// Dump2/Load2: one file per rank, ./BIN/dump-<label>-<rank %05d>-<step %010d>.bin,
//   header int32 {nx, ny, nz, vc, ne = 1, nb}, then every block's ghosted
//   (n+2vc)^3 array in block-ID order.
// Dump/Load:   one file per block, ./BIN/dump-<label>-<rank>-<id %05d>-<step>.bin,
//   header int32 {nx, ny, nz, vc, ne = 1}, then that block's array.
// A load whose file holds blocks of half the edge (n/2) fills each fine
// interior cell with its coarse cell, sf3d_copy_x2 (bsf3d.f90:18-46); the
// fine ghosts keep their values.  Any other header is an error (the
// reference calls Exit(0)).
//
// The field stays in HBM: the body streams through two pinned staging
// buffers, so the PCIe copy of one chunk overlaps the file write (or read)
// of the other.
#include <sys/stat.h>

#include <cerrno>
#include <cstdio>
#include <cstring>
#include <sstream>
#include <string>
#include <vector>

#include "bcm_internal.h"

namespace bcm {
namespace {

constexpr size_t kStage = size_t(32) << 20;  // bytes per staging buffer

std::string join(const char* dir, const char* label, int rank, int id, int step) {
  char tail[96];
  if (id < 0)
    std::snprintf(tail, sizeof tail, "-%05d-%010d.bin", rank, step);
  else
    std::snprintf(tail, sizeof tail, "-%05d-%05d-%010d.bin", rank, id, step);
  return std::string(dir) + "/dump-" + label + tail;
}

// Two pinned buffers and their completion events.  Copies still queued
// on the stream when a transfer stops early finish before the buffers go.
struct Stage {
  hipStream_t stream;
  char* buf[2] = {nullptr, nullptr};
  hipEvent_t ev[2] = {nullptr, nullptr};
  explicit Stage(hipStream_t s) : stream(s) {}
  ~Stage() {
    (void)hipStreamSynchronize(stream);
    for (int i = 0; i < 2; ++i) {
      if (ev[i]) (void)hipEventDestroy(ev[i]);
      if (buf[i]) (void)hipHostFree(buf[i]);
    }
  }
  int init() {
    for (int i = 0; i < 2; ++i) {
      if (hipHostMalloc(reinterpret_cast<void**>(&buf[i]), kStage) != hipSuccess) return BCM_ERR_NOMEM;
      if (hipEventCreateWithFlags(&ev[i], hipEventDisableTiming) != hipSuccess) return BCM_ERR_HIP;
    }
    return BCM_OK;
  }
};

struct File {
  FILE* fp = nullptr;
  ~File() {
    if (fp) std::fclose(fp);
  }
};

// Device bytes [src, src+len) appended to fp: chunk c is copied while
// chunk c-1 is written.
int stream_out(bcm_mesh* m, const char* src, size_t len, FILE* fp, Stage& st) {
  const size_t nchunk = (len + kStage - 1) / kStage;
  for (size_t c = 0; c <= nchunk; ++c) {
    if (c < nchunk) {
      const size_t off = c * kStage, n = std::min(kStage, len - off);
      BCM_HIP_TRY(hipMemcpyAsync(st.buf[c & 1], src + off, n, hipMemcpyDeviceToHost, m->stream));
      BCM_HIP_TRY(hipEventRecord(st.ev[c & 1], m->stream));
    }
    if (c > 0) {
      const size_t p = c - 1, off = p * kStage, n = std::min(kStage, len - off);
      BCM_HIP_TRY(hipEventSynchronize(st.ev[p & 1]));
      if (std::fwrite(st.buf[p & 1], 1, n, fp) != n) return BCM_ERR_IO;
    }
  }
  return BCM_OK;
}

// len bytes of fp into device memory at dst: chunk c is read while chunk
// c-1 is copied up.
int stream_in(bcm_mesh* m, char* dst, size_t len, FILE* fp, Stage& st) {
  const size_t nchunk = (len + kStage - 1) / kStage;
  for (size_t c = 0; c < nchunk; ++c) {
    const size_t off = c * kStage, n = std::min(kStage, len - off);
    // The buffer's previous upload (chunk c-2) must be done before reuse.
    if (c >= 2) BCM_HIP_TRY(hipEventSynchronize(st.ev[c & 1]));
    if (std::fread(st.buf[c & 1], 1, n, fp) != n) return BCM_ERR_IO;
    BCM_HIP_TRY(hipMemcpyAsync(dst + off, st.buf[c & 1], n, hipMemcpyHostToDevice, m->stream));
    BCM_HIP_TRY(hipEventRecord(st.ev[c & 1], m->stream));
  }
  BCM_HIP_TRY(hipStreamSynchronize(m->stream));
  return BCM_OK;
}

// sf3d_copy_x2: fine interior cell (i, j, k) of block b takes coarse cell
// (i/2, j/2, k/2) of the half-size block b; one thread per fine cell.
template <typename T>
__global__ __launch_bounds__(256) void k_copy_x2(T* __restrict__ fine, const T* __restrict__ coarse, int nblk,
                                                 int n, int vc) {
  const long long per = (long long)n * n * n;
  const long long c = (long long)blockIdx.x * 256 + threadIdx.x;
  if (c >= per * nblk) return;
  const int b = int(c / per);
  const int r = int(c % per);
  const int i = r % n, j = (r / n) % n, k = r / (n * n);
  const int nf = n + 2 * vc, h = n / 2, nc = h + 2 * vc;
  const long long fo = (long long)b * nf * nf * nf + (i + vc) + (long long)nf * ((j + vc) + (long long)nf * (k + vc));
  const long long co =
      (long long)b * nc * nc * nc + (i / 2 + vc) + (long long)nc * ((j / 2 + vc) + (long long)nc * (k / 2 + vc));
  fine[fo] = coarse[co];
}

int read_header(FILE* fp, int* h, int count) {
  return std::fread(h, sizeof(int32_t), size_t(count), fp) == size_t(count) ? BCM_OK : BCM_ERR_IO;
}

// Body of nblk blocks stored with edge nfile (= n or n/2) into blocks
// first.. of f: streamed straight in, or through a device staging area
// and k_copy_x2.
int load_blocks(bcm_field* f, int first, int nblk, int nfile, FILE* fp, Stage& st) {
  bcm_mesh* m = f->mesh;
  const size_t es = f->dtype == BCM_F64 ? 8 : 4;
  char* dev = static_cast<char*>(f->data) + size_t(first) * m->bvol * es;
  if (nfile == m->n) return stream_in(m, dev, size_t(nblk) * m->bvol * es, fp, st);
  const size_t cvol = size_t(nfile + 2 * m->vc) * (nfile + 2 * m->vc) * (nfile + 2 * m->vc);
  const int per_pass = int(std::max<size_t>(1, (size_t(4) * kStage) / (cvol * es)));
  void* tmp = nullptr;
  BCM_TRY(dev_alloc(&tmp, size_t(std::min(per_pass, nblk)) * cvol * es));
  int rc = BCM_OK;
  for (int b0 = 0; b0 < nblk && rc == BCM_OK; b0 += per_pass) {
    const int nb = std::min(per_pass, nblk - b0);
    rc = stream_in(m, static_cast<char*>(tmp), size_t(nb) * cvol * es, fp, st);
    if (rc != BCM_OK) break;
    const long long cells = (long long)nb * m->n * m->n * m->n;
    const dim3 grid(unsigned((cells + 255) / 256));
    char* out = dev + size_t(b0) * m->bvol * es;
    if (f->dtype == BCM_F64)
      hipLaunchKernelGGL(k_copy_x2<double>, grid, dim3(256), 0, m->stream, reinterpret_cast<double*>(out),
                         static_cast<const double*>(tmp), nb, m->n, m->vc);
    else
      hipLaunchKernelGGL(k_copy_x2<float>, grid, dim3(256), 0, m->stream, reinterpret_cast<float*>(out),
                         static_cast<const float*>(tmp), nb, m->n, m->vc);
    if (hipGetLastError() != hipSuccess) rc = BCM_ERR_HIP;
    // The next pass overwrites tmp: wait for the kernel.
    if (rc == BCM_OK && hipStreamSynchronize(m->stream) != hipSuccess) rc = BCM_ERR_HIP;
  }
  (void)hipStreamSynchronize(m->stream);
  dev_free(tmp);
  return rc;
}

// VtkWriter::printVTIC: the interior of each block as FP32, x fastest,
// n^3 values per block in block order.
template <typename T>
__global__ __launch_bounds__(256) void k_vtk_pack(const T* __restrict__ field, float* __restrict__ out, int first,
                                                  int nblk, int n, int vc) {
  const long long per = (long long)n * n * n;
  const long long c = (long long)blockIdx.x * 256 + threadIdx.x;
  if (c >= per * nblk) return;
  const int b = int(c / per), r = int(c % per);
  const int i = r % n, j = (r / n) % n, k = r / (n * n);
  const int nx = n + 2 * vc;
  const long long o = (long long)(first + b) * nx * nx * nx + (i + vc) + (long long)nx * ((j + vc) + (long long)nx * (k + vc));
  out[c] = float(field[o]);
}

// The XML of one block's .vti, up to the appended-data marker, as
// VtkWriter::printVTIC writes it (examples/SolverTest/VtkWriter.h:118-168).
std::string vti_head(int n, double ox, double oy, double oz, double dx, const char* label) {
  std::ostringstream o;
  o << "<VTKFile type=\"ImageData\" version=\"0.1\" byte_order=\"LittleEndian\">" << std::endl;
  o << "<ImageData WholeExtent=\"0 " << n << " 0 " << n << " 0 " << n << "\" ";
  o << "Origin=\"";
  o.setf(std::ios::scientific, std::ios::floatfield);
  o.precision(16);
  o << ox << " " << oy << " " << oz << "\" ";
  o << "Spacing=\"" << dx << " " << dx << " " << dx << "\">" << std::endl;
  o << "<Piece Extent=\"0 " << n << " 0 " << n << " 0 " << n << "\">" << std::endl;
  o << "<PointData>" << std::endl << "</PointData>" << std::endl << "<CellData>" << std::endl;
  o << "<DataArray type=\"Float32\" Name=\"" << label << "\" format=\"appended\" offset=\"" << 0 << "\"/>"
    << std::endl;
  o << "</CellData>" << std::endl << "<Coordinates>" << std::endl << "</Coordinates>" << std::endl;
  o << "</Piece>" << std::endl << "</ImageData>" << std::endl << "<AppendedData encoding=\"raw\">" << std::endl;
  o << "_";
  return o.str();
}

std::string step_dir(const char* dir, int step) {
  char buf[16];
  std::snprintf(buf, sizeof buf, "%010d", step);
  return std::string(dir) + "/" + buf;
}

int write_vti(const std::string& path, const std::string& head, const float* data, int n) {
  File out;
  out.fp = std::fopen(path.c_str(), "wb");
  if (!out.fp) return BCM_ERR_IO;
  const int32_t nbytes = int32_t(sizeof(float)) * n * n * n;
  static const char tail[] = "\n</AppendedData>\n</VTKFile>\n";
  if (std::fwrite(head.data(), 1, head.size(), out.fp) != head.size()) return BCM_ERR_IO;
  if (std::fwrite(&nbytes, sizeof nbytes, 1, out.fp) != 1) return BCM_ERR_IO;
  if (std::fwrite(data, 1, size_t(nbytes), out.fp) != size_t(nbytes)) return BCM_ERR_IO;
  if (std::fwrite(tail, 1, sizeof tail - 1, out.fp) != sizeof tail - 1) return BCM_ERR_IO;
  const int rc = std::fclose(out.fp);
  out.fp = nullptr;
  return rc == 0 ? BCM_OK : BCM_ERR_IO;
}

int check_args(bcm_field* f, const char* dir, const char* label) {
  if (!f || !dir || !label) return BCM_ERR_ARG;
  return BCM_OK;
}

}  // namespace

int dump2_raw(bcm_mesh* m, const void* dev, int nblk, int esize, const char* dir, const char* label, int rank,
              int step) {
  if (!m || !dev || nblk < 0 || !dir || !label) return BCM_ERR_ARG;
  if (::mkdir(dir, 0755) != 0 && errno != EEXIST) return BCM_ERR_IO;
  Stage st(m->stream);
  BCM_TRY(st.init());
  File out;
  out.fp = std::fopen(join(dir, label, rank, -1, step).c_str(), "wb");
  if (!out.fp) return BCM_ERR_IO;
  const int32_t h[6] = {m->n, m->n, m->n, m->vc, 1, nblk};
  if (std::fwrite(h, sizeof h, 1, out.fp) != 1) return BCM_ERR_IO;
  BCM_HIP_TRY(hipStreamSynchronize(m->stream));  // the data's producers
  BCM_TRY(stream_out(m, static_cast<const char*>(dev), size_t(nblk) * m->bvol * size_t(esize), out.fp, st));
  if (std::fclose(out.fp) != 0) {
    out.fp = nullptr;
    return BCM_ERR_IO;
  }
  out.fp = nullptr;
  return BCM_OK;
}

int load2_header(const char* dir, const char* label, int rank, int step, int n, int vc, int nblk, int* nfile) {
  if (!dir || !label) return BCM_ERR_ARG;
  File in;
  in.fp = std::fopen(join(dir, label, rank, -1, step).c_str(), "rb");
  if (!in.fp) return BCM_ERR_IO;
  int32_t h[6];
  BCM_TRY(read_header(in.fp, h, 6));
  const bool same = h[0] == n && h[1] == n && h[2] == n;
  const bool half = 2 * h[0] == n && 2 * h[1] == n && 2 * h[2] == n;
  if (!(same || half) || h[3] != vc || h[4] != 1 || h[5] != nblk) return BCM_ERR_ARG;
  if (nfile) *nfile = h[0];
  return BCM_OK;
}

int load2_into(bcm_field* f, int first, int nblk, const char* dir, const char* label, int rank, int step) {
  bcm_mesh* m = f->mesh;
  if (first < 0 || nblk < 0 || first + nblk > m->nblocks) return BCM_ERR_ARG;
  int nfile = 0;
  BCM_TRY(load2_header(dir, label, rank, step, m->n, m->vc, nblk, &nfile));
  File in;
  in.fp = std::fopen(join(dir, label, rank, -1, step).c_str(), "rb");
  if (!in.fp) return BCM_ERR_IO;
  int32_t h[6];
  BCM_TRY(read_header(in.fp, h, 6));
  Stage st(m->stream);
  BCM_TRY(st.init());
  BCM_HIP_TRY(hipStreamSynchronize(m->stream));
  return load_blocks(f, first, nblk, nfile, in.fp, st);
}

}  // namespace bcm

using namespace bcm;

extern "C" int bcm_field_dump2(bcm_field_t f, const char* dir, const char* label, int rank, int step) {
  BCM_TRY(check_args(f, dir, label));
  return dump2_raw(f->mesh, f->data, f->mesh->nblocks, f->dtype == BCM_F64 ? 8 : 4, dir, label, rank, step);
}

extern "C" int bcm_field_load2(bcm_field_t f, const char* dir, const char* label, int rank, int step) {
  BCM_TRY(check_args(f, dir, label));
  return load2_into(f, 0, f->mesh->nblocks, dir, label, rank, step);
}

extern "C" int bcm_field_dump(bcm_field_t f, const char* dir, const char* label, int rank, int step) {
  BCM_TRY(check_args(f, dir, label));
  bcm_mesh* m = f->mesh;
  if (::mkdir(dir, 0755) != 0 && errno != EEXIST) return BCM_ERR_IO;
  Stage st(m->stream);
  BCM_TRY(st.init());
  BCM_HIP_TRY(hipStreamSynchronize(m->stream));
  const size_t bbytes = f->bytes / size_t(m->nblocks);
  const int32_t h[5] = {m->n, m->n, m->n, m->vc, 1};
  for (int id = 0; id < m->nblocks; ++id) {
    File out;
    out.fp = std::fopen(join(dir, label, rank, id, step).c_str(), "wb");
    if (!out.fp) return BCM_ERR_IO;
    if (std::fwrite(h, sizeof h, 1, out.fp) != 1) return BCM_ERR_IO;
    BCM_TRY(stream_out(m, static_cast<const char*>(f->data) + size_t(id) * bbytes, bbytes, out.fp, st));
    const int rc = std::fclose(out.fp);
    out.fp = nullptr;
    if (rc != 0) return BCM_ERR_IO;
  }
  return BCM_OK;
}

extern "C" int bcm_field_load(bcm_field_t f, const char* dir, const char* label, int rank, int step) {
  BCM_TRY(check_args(f, dir, label));
  bcm_mesh* m = f->mesh;
  Stage st(m->stream);
  BCM_TRY(st.init());
  BCM_HIP_TRY(hipStreamSynchronize(m->stream));
  for (int id = 0; id < m->nblocks; ++id) {
    File in;
    in.fp = std::fopen(join(dir, label, rank, id, step).c_str(), "rb");
    if (!in.fp) return BCM_ERR_IO;
    int32_t h[5];
    BCM_TRY(read_header(in.fp, h, 5));
    const bool same = h[0] == m->n && h[1] == m->n && h[2] == m->n;
    const bool half = 2 * h[0] == m->n && 2 * h[1] == m->n && 2 * h[2] == m->n;
    if (!(same || half) || h[3] != m->vc || h[4] != 1) return BCM_ERR_ARG;
    BCM_TRY(load_blocks(f, id, 1, h[0], in.fp, st));
  }
  return BCM_OK;
}

extern "C" int bcm_field_write_vtk(bcm_field_t f, const char* dir, const char* name, int step, int rank,
                                   int start_id, int difflevel, const double root_origin[3], double root_length) {
  BCM_TRY(check_args(f, dir, name));
  if (!root_origin || !(root_length > 0.0) || difflevel < 0 || start_id < 0) return BCM_ERR_ARG;
  bcm_mesh* m = f->mesh;
  const int n = m->n;
  if (::mkdir(dir, 0755) != 0 && errno != EEXIST) return BCM_ERR_IO;
  const std::string sdir = step_dir(dir, step);
  if (::mkdir(sdir.c_str(), 0755) != 0 && errno != EEXIST) return BCM_ERR_IO;
  Stage st(m->stream);
  BCM_TRY(st.init());
  const size_t per = size_t(n) * n * n;
  const int group = int(std::max<size_t>(1, kStage / (per * sizeof(float))));
  void* tmp = nullptr;
  BCM_TRY(dev_alloc(&tmp, size_t(std::min(group, m->nblocks)) * per * sizeof(float)));
  int rc = BCM_OK;
  BCM_HIP_TRY(hipStreamSynchronize(m->stream));
  char fname[256];
  // Group g is packed and copied down while group g-1 is written out.
  const int ngroups = (m->nblocks + group - 1) / group;
  for (int g = 0; g <= ngroups && rc == BCM_OK; ++g) {
    if (g < ngroups) {
      const int b0 = g * group, nb = std::min(group, m->nblocks - b0);
      // The staging buffer of group g-2 was written out before this point.
      const long long cells = (long long)nb * (long long)per;
      const dim3 grid(unsigned((cells + 255) / 256));
      if (f->dtype == BCM_F64)
        hipLaunchKernelGGL(k_vtk_pack<double>, grid, dim3(256), 0, m->stream, static_cast<const double*>(f->data),
                           static_cast<float*>(tmp), b0, nb, n, m->vc);
      else
        hipLaunchKernelGGL(k_vtk_pack<float>, grid, dim3(256), 0, m->stream, static_cast<const float*>(f->data),
                           static_cast<float*>(tmp), b0, nb, n, m->vc);
      if (hipGetLastError() != hipSuccess ||
          hipMemcpyAsync(st.buf[g & 1], tmp, size_t(nb) * per * sizeof(float), hipMemcpyDeviceToHost, m->stream) !=
              hipSuccess ||
          hipEventRecord(st.ev[g & 1], m->stream) != hipSuccess) {
        rc = BCM_ERR_HIP;
        break;
      }
    }
    if (g > 0) {
      const int p = g - 1, b0 = p * group, nb = std::min(group, m->nblocks - b0);
      if (hipEventSynchronize(st.ev[p & 1]) != hipSuccess) {
        rc = BCM_ERR_HIP;
        break;
      }
      const float* data = reinterpret_cast<const float*>(st.buf[p & 1]);
      for (int b = 0; b < nb && rc == BCM_OK; ++b) {
        const bcm_block_info& bi = m->blocks[size_t(b0 + b)];
        const double dx = bi.block_size[0] / n;
        std::snprintf(fname, sizeof fname, "/data-%s-%05d-%05d-%010d.vti", name, rank, b0 + b, step);
        rc = write_vti(sdir + fname, vti_head(n, bi.origin[0], bi.origin[1], bi.origin[2], dx, name),
                       data + size_t(b) * per, n);
      }
    }
    // The next pack reuses tmp: its copy into the staging buffer must be done.
    if (rc == BCM_OK && g < ngroups && hipEventSynchronize(st.ev[g & 1]) != hipSuccess) rc = BCM_ERR_HIP;
  }
  (void)hipStreamSynchronize(m->stream);
  dev_free(tmp);
  BCM_TRY(rc);
  // The hierarchical-box index (VtkWriter::writeScalar, VtkWriter.h:771-889):
  // one DataSet per block with its level and its box in cells of that level.
  if (rank != 0) return BCM_OK;
  File out;
  std::ostringstream o;
  o << "<VTKFile type=\"vtkHierarchicalBoxDataSet\" version=\"0.1\">" << std::endl;
  o << "<vtkHierarchicalBoxDataSet>" << std::endl;
  for (int l = 0; l < difflevel + 1; ++l) o << "<RefinementRatio level=\"" << l << "\" refinement=\"2\"/>" << std::endl;
  o << std::endl;
  for (int id = 0; id < m->nblocks; ++id) {
    const bcm_block_info& bi = m->blocks[size_t(id)];
    const int level = bi.level;
    const double lx = double(n) * double(1 << level);
    double box[3][2];
    for (int a = 0; a < 3; ++a) {
      const double o2 = (bi.origin[a] - root_origin[a]) / root_length;  // origin in root lengths
      box[a][0] = o2 * lx;
      box[a][1] = box[a][0] + n - 1;
    }
    std::snprintf(fname, sizeof fname, "./%010d/data-%s-%05d-%05d-%010d.vti", step, name, rank, id, step);
    o << "<DataSet group=\"" << level << "\" dataset=\"" << id + start_id << "\" amr_box=\"" << box[0][0] << " "
      << box[0][1] << " " << box[1][0] << " " << box[1][1] << " " << box[2][0] << " " << box[2][1] << "\" file=\""
      << fname << "\"/>" << std::endl;
  }
  o << "</vtkHierarchicalBoxDataSet>" << std::endl << "</VTKFile>" << std::endl;
  const std::string text = o.str();
  char iname[256];
  std::snprintf(iname, sizeof iname, "/data-%s-%010d.vthb", name, step);
  out.fp = std::fopen((std::string(dir) + iname).c_str(), "wb");
  if (!out.fp) return BCM_ERR_IO;
  if (std::fwrite(text.data(), 1, text.size(), out.fp) != text.size()) return BCM_ERR_IO;
  const int frc = std::fclose(out.fp);
  out.fp = nullptr;
  return frc == 0 ? BCM_OK : BCM_ERR_IO;
}
