// System files: the linear system as one dense array of shape (nz, ny, nx, 8)
// of little-endian doubles, the dataset the reference's linear test reads and
// writes with --system_in / --system_out (src/test/linear/main.cpp:181-189,
// HDF5 dataset "data"), stored here in the NumPy .npy format (version 1.0,
// C order): a 10-byte preamble, a Python-literal header padded with spaces
// to a multiple of 64 bytes and ending in '\n', then the raw data.  numpy.load
// reads these files as they are (no pickles), and any .npy of that dtype and
// shape written by numpy.save is accepted.
#include <cstdio>
#include <cstring>
#include <string>

#include "../../include/aphpois.h"

namespace {

thread_local std::string g_err;

int fail(int code, const std::string& msg) {
  g_err = msg;
  return code;
}

// Parses the header dictionary; returns false if the array is not a C-order
// little-endian float64 array of 4 dimensions with last dimension 8.
bool parse_header(const std::string& h, long long shape[4]) {
  auto find_value = [&](const char* key) -> std::string {
    const std::string k = std::string("'") + key + "'";
    size_t p = h.find(k);
    if (p == std::string::npos) return "";
    p = h.find(':', p + k.size());
    if (p == std::string::npos) return "";
    ++p;
    while (p < h.size() && h[p] == ' ') ++p;
    if (p >= h.size()) return "";
    if (h[p] == '(') {
      const size_t e = h.find(')', p);
      return e == std::string::npos ? "" : h.substr(p, e - p + 1);
    }
    size_t e = p;
    while (e < h.size() && h[e] != ',' && h[e] != '}') ++e;
    return h.substr(p, e - p);
  };
  const std::string descr = find_value("descr");
  if (descr != "'<f8'") return false;
  const std::string fo = find_value("fortran_order");
  if (fo.compare(0, 5, "False") != 0) return false;
  const std::string sh = find_value("shape");
  if (sh.size() < 2) return false;
  int k = 0;
  size_t p = 1;
  while (p < sh.size() && k < 5) {
    while (p < sh.size() && (sh[p] == ' ' || sh[p] == ',')) ++p;
    if (p >= sh.size() || sh[p] == ')') break;
    long long v = 0;
    bool any = false;
    while (p < sh.size() && sh[p] >= '0' && sh[p] <= '9') {
      v = v * 10 + (sh[p] - '0');
      ++p;
      any = true;
    }
    if (!any) return false;
    if (k < 4) shape[k] = v;
    ++k;
  }
  return k == 4 && shape[3] == 8 && shape[0] > 0 && shape[1] > 0 && shape[2] > 0;
}

int read_header(std::FILE* f, long long shape[4], long* data_offset) {
  unsigned char pre[10];
  if (std::fread(pre, 1, 10, f) != 10 || std::memcmp(pre, "\x93NUMPY", 6) != 0)
    return fail(APHPOIS_EINVAL, "not an .npy file");
  long hlen = 0;
  long off = 10;
  if (pre[6] == 1) {
    hlen = pre[8] | (pre[9] << 8);
  } else if (pre[6] == 2 || pre[6] == 3) {
    unsigned char b[2];
    if (std::fread(b, 1, 2, f) != 2) return fail(APHPOIS_EINVAL, "truncated .npy header");
    hlen = (long)pre[8] | ((long)pre[9] << 8) | ((long)b[0] << 16) | ((long)b[1] << 24);
    off = 12;
  } else {
    return fail(APHPOIS_EINVAL, "unsupported .npy version");
  }
  if (hlen <= 0 || hlen > (1L << 20)) return fail(APHPOIS_EINVAL, "bad .npy header length");
  std::string h((size_t)hlen, '\0');
  if (std::fread(&h[0], 1, (size_t)hlen, f) != (size_t)hlen)
    return fail(APHPOIS_EINVAL, "truncated .npy header");
  if (!parse_header(h, shape))
    return fail(APHPOIS_EINVAL,
                "the file must hold a C-order float64 array of shape (nz, ny, nx, 8)");
  *data_offset = off + hlen;
  return APHPOIS_OK;
}

}  // namespace

extern "C" {

const char* aphpois_system_file_error(void) { return g_err.c_str(); }

int aphpois_system_file_info(const char* path, int* nx, int* ny, int* nz) {
  if (!path) return fail(APHPOIS_EINVAL, "path is NULL");
  std::FILE* f = std::fopen(path, "rb");
  if (!f) return fail(APHPOIS_EINVAL, std::string("cannot open ") + path);
  long long shape[4];
  long off = 0;
  const int rc = read_header(f, shape, &off);
  std::fclose(f);
  if (rc != APHPOIS_OK) return rc;
  if (shape[0] > 0x7fffffff || shape[1] > 0x7fffffff || shape[2] > 0x7fffffff)
    return fail(APHPOIS_EINVAL, "grid too large");
  if (nz) *nz = (int)shape[0];
  if (ny) *ny = (int)shape[1];
  if (nx) *nx = (int)shape[2];
  return APHPOIS_OK;
}

int aphpois_system_file_read(const char* path, double* expr, long long count) {
  if (!path || !expr) return fail(APHPOIS_EINVAL, "path or expr is NULL");
  std::FILE* f = std::fopen(path, "rb");
  if (!f) return fail(APHPOIS_EINVAL, std::string("cannot open ") + path);
  long long shape[4];
  long off = 0;
  int rc = read_header(f, shape, &off);
  if (rc == APHPOIS_OK) {
    const long long need = shape[0] * shape[1] * shape[2] * 8;
    if (need != count) {
      rc = fail(APHPOIS_EINVAL, "count does not match the file's nx*ny*nz*8");
    } else if (std::fseek(f, off, SEEK_SET) != 0 ||
               std::fread(expr, sizeof(double), (size_t)count, f) != (size_t)count) {
      rc = fail(APHPOIS_EINVAL, "truncated system file");
    }
  }
  std::fclose(f);
  return rc;
}

int aphpois_system_file_write(const char* path, int nx, int ny, int nz, const double* expr) {
  if (!path || !expr) return fail(APHPOIS_EINVAL, "path or expr is NULL");
  if (nx < 1 || ny < 1 || nz < 1) return fail(APHPOIS_EINVAL, "grid size must be >= 1");
  std::string h = "{'descr': '<f8', 'fortran_order': False, 'shape': (" + std::to_string(nz) +
                  ", " + std::to_string(ny) + ", " + std::to_string(nx) + ", 8), }";
  // pad so that preamble + header is a multiple of 64 bytes, ending in '\n'
  const size_t total = 10 + h.size() + 1;
  h.append((64 - total % 64) % 64, ' ');
  h.push_back('\n');
  if (h.size() > 65535) return fail(APHPOIS_EINVAL, "header too long");
  std::FILE* f = std::fopen(path, "wb");
  if (!f) return fail(APHPOIS_EINVAL, std::string("cannot create ") + path);
  unsigned char pre[10] = {0x93, 'N', 'U', 'M', 'P', 'Y', 1, 0,
                           (unsigned char)(h.size() & 0xff), (unsigned char)(h.size() >> 8)};
  const size_t count = (size_t)nx * ny * nz * 8;
  bool ok = std::fwrite(pre, 1, 10, f) == 10 && std::fwrite(h.data(), 1, h.size(), f) == h.size() &&
            std::fwrite(expr, sizeof(double), count, f) == count;
  ok = (std::fclose(f) == 0) && ok;
  return ok ? APHPOIS_OK : fail(APHPOIS_EINVAL, std::string("cannot write ") + path);
}

}  // extern "C"
