// ms_output.cpp -- VTU output of the cell fields, host only.
//
// Restates vtk_write (mod_uns_output.f90:81-240, default tetrahedral fan)
// with vtu_write's two layouts (mod_vtu_writer.f90:159-321 inline ASCII,
// :329-455 appended raw binary), so a file written here is byte-identical to
// the reference's for the same mesh and fields and the reference's case
// scripts read it unchanged:
//   points: mesh nodes, then one centroid point per cell;
//   cells:  every face of cell c (its faces in ascending face order) fanned
//           from its first node, (n1, nj, nj+1) when c owns the face and
//           (n1, nj+1, nj) otherwise, closed by the centroid point: VTK_TETRA;
//   cell data: pressure, velocity, temperature, optionally temperature_solid
//           and heat_source_solid, each replicated on the cell's tetrahedra.
// ASCII numbers use Fortran's ES24.16E3 edit descriptor (17 significant
// digits, three-digit exponent, right-justified in 24 columns, one blank
// before each); integers one per line.
//
// Also the field dump v1 of write_field_dump / read_field_dump
// (mod_uns_restart.f90:77-152, :166-334): a '#' header (version, source
// mesh, ncells, ts_order, generation time, field list), then u, p, T,
// u_old, u_old_old, T_old, T_old_old one cell per line in ES24.16 (vectors
// '(3(es24.16,1x))'), each block closed by a '# end <name>' line; read back
// list-directed with the reference's error codes.
#include <cctype>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <ctime>
#include <memory>
#include <string>
#include <vector>

#include "ms_common.hpp"

namespace ms {
namespace {

// Fortran ES24.16E3 of v into out (24 characters, no terminator needed)
int es24(double v, char *out) {
  char b[64];
  if (std::isnan(v)) {
    std::snprintf(b, sizeof b, "NaN");
  } else if (std::isinf(v)) {
    std::snprintf(b, sizeof b, v < 0 ? "-Infinity" : "Infinity");
  } else {
    char m[64];
    std::snprintf(m, sizeof m, "%.16E", v);
    char *e = std::strchr(m, 'E');
    const int ex = std::atoi(e + 1);
    *e = '\0';
    std::snprintf(b, sizeof b, "%sE%c%03d", m, ex < 0 ? '-' : '+', ex < 0 ? -ex : ex);
  }
  const int len = (int)std::strlen(b);
  if (len > 24) {
    std::memset(out, '*', 24);
    return 24;
  }
  std::memset(out, ' ', 24 - len);
  std::memcpy(out + 24 - len, b, len);
  return 24;
}

// Fortran ES24.16 (no exponent width): E+dd for |exponent| <= 99, +ddd
// (no letter) beyond, right-justified in 24 columns
int es24_2(double v, char *out) {
  char b[64];
  if (std::isnan(v)) {
    std::snprintf(b, sizeof b, "NaN");
  } else if (std::isinf(v)) {
    std::snprintf(b, sizeof b, v < 0 ? "-Infinity" : "Infinity");
  } else {
    char m[64];
    std::snprintf(m, sizeof m, "%.16E", v);
    char *e = std::strchr(m, 'E');
    const int ex = std::atoi(e + 1);
    *e = '\0';
    const int a = ex < 0 ? -ex : ex;
    if (a <= 99) std::snprintf(b, sizeof b, "%sE%c%02d", m, ex < 0 ? '-' : '+', a);
    else std::snprintf(b, sizeof b, "%s%c%03d", m, ex < 0 ? '-' : '+', a);
  }
  const int len = (int)std::strlen(b);
  if (len > 24) {
    std::memset(out, '*', 24);
    return 24;
  }
  std::memset(out, ' ', 24 - len);
  std::memcpy(out + 24 - len, b, len);
  return 24;
}

// one Fortran list-directed real: D/d exponents, and the letterless form
// d.ddd+ddd that ES editing writes for three-digit exponents
bool parse_real(const char *t, size_t n, double &v) {
  std::string s(t, n);
  for (char &c : s)
    if (c == 'D' || c == 'd') c = 'E';
  if (s.find_first_of("Ee") == std::string::npos) {
    const size_t k = s.find_last_of("+-");
    if (k != std::string::npos && k > 0 && std::isdigit((unsigned char)s[k - 1])) s.insert(k, "E");
  }
  char *end = nullptr;
  v = std::strtod(s.c_str(), &end);
  return end && *end == '\0' && !s.empty();
}

struct Out {
  FILE *f = nullptr;
  explicit Out(const char *path, bool bin) {
    f = std::fopen(path, bin ? "wb" : "w");
    if (!f) fail(MS_EINVAL, std::string("cannot open ") + path);
    std::setvbuf(f, nullptr, _IOFBF, 1 << 22);
  }
  ~Out() {
    if (f) std::fclose(f);
  }
  void s(const std::string &t) { std::fwrite(t.data(), 1, t.size(), f); }
  void raw(const void *p, size_t n) { std::fwrite(p, 1, n, f); }
  void reals(const double *v, int n) {   // one line: n x (1x, ES24.16E3)
    char line[25 * 8 + 2];
    int k = 0;
    for (int i = 0; i < n; i++) {
      line[k++] = ' ';
      k += es24(v[i], line + k);
    }
    line[k++] = '\n';
    std::fwrite(line, 1, k, f);
  }
  void ints(long long v) { std::fprintf(f, "%lld\n", v); }
  void dump_line(const double *v, int n) {   // '(3(es24.16,1x))' or '(es24.16)'
    char line[25 * 4 + 2];
    int k = 0;
    for (int i = 0; i < n; i++) {
      k += es24_2(v[i], line + k);
      if (n > 1) line[k++] = ' ';
    }
    line[k++] = '\n';
    std::fwrite(line, 1, k, f);
  }
  void close() {
    if (std::fclose(f) != 0) {
      f = nullptr;
      fail(MS_EINVAL, "output file write failed");
    }
    f = nullptr;
  }
};

struct Data {
  const char *name;
  int ncomp;
  const double *val;   // (ncomp, ncells)
};

uint32_t u32_head(long long nbytes) {
  if (nbytes < 0 || nbytes > 4294967295LL) fail(MS_EINVAL, "VTU block exceeds the UInt32 header range");
  return (uint32_t)nbytes;
}

}  // namespace

void vtu_write(const char *path, bool binary, int nnodes, const double *x, int nfaces, const long long *fptr,
               const int *fnodes, const int *c0, const int *c1, int ncells, const double *xc, const double *p,
               const double *u, const double *T, const double *T_s, const double *q_src) {
  if (!path || nnodes <= 0 || nfaces <= 0 || ncells <= 0 || !x || !fptr || !fnodes || !c0 || !c1 || !xc || !p ||
      !u || !T)
    fail(MS_EINVAL, "ms_vtu_write: empty mesh or null argument");
  // cell -> faces, ascending face order (build_connectivity's fill order)
  std::vector<int> cptr(ncells + 1, 0), cf;
  for (int f = 0; f < nfaces; f++) {
    if (c0[f] < 1 || c0[f] > ncells || c1[f] < 0 || c1[f] > ncells) fail(MS_EINVAL, "ms_vtu_write: bad face cells");
    cptr[c0[f]]++;
    if (c1[f] > 0) cptr[c1[f]]++;
  }
  for (int c = 0; c < ncells; c++) cptr[c + 1] += cptr[c];
  cf.resize(cptr[ncells]);
  {
    std::vector<int> fill(cptr.begin(), cptr.end() - 1);
    for (int f = 0; f < nfaces; f++) {
      cf[fill[c0[f] - 1]++] = f;
      if (c1[f] > 0) cf[fill[c1[f] - 1]++] = f;
    }
  }
  long long ntets = 0;
  std::vector<int> nsub(ncells, 0);
  for (int c = 0; c < ncells; c++)
    for (int k = cptr[c]; k < cptr[c + 1]; k++) {
      const int nn = (int)(fptr[cf[k] + 1] - fptr[cf[k]]);
      if (nn < 3) fail(MS_EINVAL, "ms_vtu_write: face with fewer than 3 nodes");
      nsub[c] += nn - 2;
    }
  for (int c = 0; c < ncells; c++) ntets += nsub[c];
  // connectivity (0-based point ids)
  std::vector<int32_t> conn(4 * (size_t)ntets);
  size_t it = 0;
  for (int c = 0; c < ncells; c++) {
    const int cen = nnodes + c;
    for (int k = cptr[c]; k < cptr[c + 1]; k++) {
      const int f = cf[k];
      const int *nd = fnodes + fptr[f];
      const int nn = (int)(fptr[f + 1] - fptr[f]);
      const bool own = c0[f] == c + 1;
      for (int j = 1; j < nn - 1; j++) {
        conn[it++] = nd[0] - 1;
        conn[it++] = (own ? nd[j] : nd[j + 1]) - 1;
        conn[it++] = (own ? nd[j + 1] : nd[j]) - 1;
        conn[it++] = cen;
      }
    }
  }
  const long long npts = (long long)nnodes + ncells;
  std::vector<Data> cd = {{"pressure", 1, p}, {"velocity", 3, u}, {"temperature", 1, T}};
  if (T_s) cd.push_back({"temperature_solid", 1, T_s});
  if (q_src) cd.push_back({"heat_source_solid", 1, q_src});
  auto point = [&](long long i) { return i < nnodes ? x + 3 * i : xc + 3 * (i - nnodes); };

  Out o(path, binary);
  o.s("<?xml version=\"1.0\"?>\n");
  if (!binary) {
    o.s("<VTKFile type=\"UnstructuredGrid\" version=\"1.0\" byte_order=\"LittleEndian\">\n");
    o.s("  <UnstructuredGrid>\n");
    o.s("    <Piece NumberOfPoints=\"" + std::to_string(npts) + "\" NumberOfCells=\"" + std::to_string(ntets) +
        "\">\n");
    o.s("      <Points>\n");
    o.s("        <DataArray type=\"Float64\" Name=\"Coordinates\" NumberOfComponents=\"3\" format=\"ascii\">\n");
    for (long long i = 0; i < npts; i++) o.reals(point(i), 3);
    o.s("        </DataArray>\n");
    o.s("      </Points>\n");
    o.s("      <Cells>\n");
    o.s("        <DataArray type=\"Int32\" Name=\"connectivity\" format=\"ascii\">\n");
    for (size_t i = 0; i < conn.size(); i++) o.ints(conn[i]);
    o.s("        </DataArray>\n");
    o.s("        <DataArray type=\"Int32\" Name=\"offsets\" format=\"ascii\">\n");
    for (long long i = 1; i <= ntets; i++) o.ints(4 * i);
    o.s("        </DataArray>\n");
    o.s("        <DataArray type=\"UInt8\" Name=\"types\" format=\"ascii\">\n");
    for (long long i = 0; i < ntets; i++) o.ints(10);
    o.s("        </DataArray>\n");
    o.s("      </Cells>\n");
    o.s("      <CellData>\n");
    for (const Data &d : cd) {
      o.s("        <DataArray type=\"Float64\" Name=\"" + std::string(d.name) + "\" NumberOfComponents=\"" +
          std::to_string(d.ncomp) + "\" format=\"ascii\">\n");
      for (int c = 0; c < ncells; c++)
        for (int j = 0; j < nsub[c]; j++) o.reals(d.val + (size_t)d.ncomp * c, d.ncomp);
      o.s("        </DataArray>\n");
    }
    o.s("      </CellData>\n");
    o.s("    </Piece>\n");
    o.s("  </UnstructuredGrid>\n");
    o.s("</VTKFile>\n");
    o.close();
    return;
  }
  // appended raw binary: Float32 payload, UInt32 block headers, offsets from
  // the first byte after '_'
  const long long off_pts = 0, off_conn = off_pts + 4 + npts * 12, off_offs = off_conn + 4 + 4 * (long long)conn.size();
  const long long off_typs = off_offs + 4 + ntets * 4;
  std::vector<long long> off_cd(cd.size());
  long long cur = off_typs + 4 + ntets;
  for (size_t j = 0; j < cd.size(); j++) {
    off_cd[j] = cur;
    cur += 4 + ntets * cd[j].ncomp * 4;
  }
  o.s("<VTKFile type=\"UnstructuredGrid\" version=\"1.0\" byte_order=\"LittleEndian\" header_type=\"UInt32\">\n");
  o.s("  <UnstructuredGrid>\n");
  o.s("    <Piece NumberOfPoints=\"" + std::to_string(npts) + "\" NumberOfCells=\"" + std::to_string(ntets) + "\">\n");
  o.s("      <Points>\n");
  o.s("        <DataArray type=\"Float32\" Name=\"Coordinates\" NumberOfComponents=\"3\" format=\"appended\" "
      "offset=\"" + std::to_string(off_pts) + "\"/>\n");
  o.s("      </Points>\n");
  o.s("      <Cells>\n");
  o.s("        <DataArray type=\"Int32\" Name=\"connectivity\" format=\"appended\" offset=\"" +
      std::to_string(off_conn) + "\"/>\n");
  o.s("        <DataArray type=\"Int32\" Name=\"offsets\" format=\"appended\" offset=\"" + std::to_string(off_offs) +
      "\"/>\n");
  o.s("        <DataArray type=\"UInt8\" Name=\"types\" format=\"appended\" offset=\"" + std::to_string(off_typs) +
      "\"/>\n");
  o.s("      </Cells>\n");
  o.s("      <CellData>\n");
  for (size_t j = 0; j < cd.size(); j++)
    o.s("        <DataArray type=\"Float32\" Name=\"" + std::string(cd[j].name) + "\" NumberOfComponents=\"" +
        std::to_string(cd[j].ncomp) + "\" format=\"appended\" offset=\"" + std::to_string(off_cd[j]) + "\"/>\n");
  o.s("      </CellData>\n");
  o.s("    </Piece>\n");
  o.s("  </UnstructuredGrid>\n");
  o.s("  <AppendedData encoding=\"raw\">\n_");
  uint32_t h = u32_head(npts * 12);
  o.raw(&h, 4);
  {
    std::vector<float> v(3 * (size_t)npts);
    for (long long i = 0; i < npts; i++)
      for (int d = 0; d < 3; d++) v[3 * i + d] = (float)point(i)[d];
    o.raw(v.data(), v.size() * 4);
  }
  h = u32_head(4 * (long long)conn.size());
  o.raw(&h, 4);
  o.raw(conn.data(), conn.size() * 4);
  {
    h = u32_head(ntets * 4);
    o.raw(&h, 4);
    std::vector<int32_t> offs(ntets);
    for (long long i = 0; i < ntets; i++) offs[i] = (int32_t)(4 * (i + 1));
    o.raw(offs.data(), offs.size() * 4);
    h = u32_head(ntets);
    o.raw(&h, 4);
    std::vector<unsigned char> ty(ntets, 10);
    o.raw(ty.data(), ty.size());
  }
  for (const Data &d : cd) {
    h = u32_head(ntets * d.ncomp * 4);
    o.raw(&h, 4);
    std::vector<float> v((size_t)ntets * d.ncomp);
    size_t q = 0;
    for (int c = 0; c < ncells; c++)
      for (int j = 0; j < nsub[c]; j++)
        for (int k = 0; k < d.ncomp; k++) v[q++] = (float)d.val[(size_t)d.ncomp * c + k];
    o.raw(v.data(), v.size() * 4);
  }
  o.s("\n  </AppendedData>\n");
  o.s("</VTKFile>\n");
  o.close();
}

// ------------------------------------------------------------ field dump v1
void dump_write(const char *path, int ncells, int ts_order, const char *source_mesh, const char *generated,
                const double *const f[7]) {
  if (!path || ncells <= 0) fail(MS_EINVAL, "ms_dump_write: null path or empty mesh");
  for (int k = 0; k < 7; k++)
    if (!f[k]) fail(MS_EINVAL, "ms_dump_write: null field");
  std::string gen;
  if (generated && generated[0]) {
    gen = generated;
  } else {
    std::time_t t = std::time(nullptr);
    std::tm tm{};
    localtime_r(&t, &tm);
    char b[32];
    std::strftime(b, sizeof b, "%Y-%m-%d %H:%M:%S", &tm);
    gen = b;
  }
  std::string src = source_mesh ? source_mesh : "";
  while (!src.empty() && src.back() == ' ') src.pop_back();
  Out o(path, false);
  o.s("# UNSolver field dump v1\n");
  o.s("# source_mesh: " + src + "\n");
  o.s("# ncells: " + std::to_string(ncells) + "\n");
  o.s("# ts_order: " + std::to_string(ts_order) + "\n");
  o.s("# generated: " + gen + "\n");
  o.s("# fields: u p T u_old u_old_old T_old T_old_old\n");
  static const char *names[7] = {"u", "p", "T", "u_old", "u_old_old", "T_old", "T_old_old"};
  static const int width[7] = {3, 1, 1, 3, 3, 1, 1};
  for (int k = 0; k < 7; k++) {
    for (int c = 0; c < ncells; c++) o.dump_line(f[k] + (size_t)width[k] * c, width[k]);
    o.s(std::string("# end ") + names[k] + "\n");
  }
  o.close();
}

int dump_read(const char *path, int ncells, double *const f[7], int *ts_order, std::string &src) {
  src.clear();
  FILE *fp = path ? std::fopen(path, "r") : nullptr;
  if (!fp) return 1;
  std::unique_ptr<FILE, int (*)(FILE *)> guard(fp, std::fclose);
  std::string line;
  auto getline = [&](std::string &l) -> bool {
    l.clear();
    int ch;
    bool any = false;
    while ((ch = std::fgetc(fp)) != EOF) {
      any = true;
      if (ch == '\n') return true;
      l.push_back((char)ch);
    }
    return any;
  };
  auto ltrim = [](const std::string &x) {
    size_t i = 0;
    while (i < x.size() && (x[i] == ' ' || x[i] == '\t')) i++;
    return x.substr(i);
  };
  auto after_colon_int = [&](const std::string &l, int &v) {
    const size_t k = l.find(':');
    v = (k == std::string::npos) ? 0 : std::atoi(l.c_str() + k + 1);
  };
  int ver = 0, n_in = 0, ts_in = 0;
  const std::string vmark = "# UNSolver field dump v";
  while (true) {
    if (!getline(line)) return 2;          // ended in the header
    line = ltrim(line);
    while (!line.empty() && (line.back() == ' ' || line.back() == '\r')) line.pop_back();
    if (line.empty()) continue;
    if (line[0] != '#') return 2;          // malformed header
    if (line.find(vmark) != std::string::npos) {
      ver = std::atoi(line.c_str() + line.find(vmark) + vmark.size());
    } else if (line.find("# source_mesh:") != std::string::npos) {
      src = ltrim(line.substr(line.find(':') + 1));
    } else if (line.find("# ncells:") != std::string::npos) {
      after_colon_int(line, n_in);
    } else if (line.find("# ts_order:") != std::string::npos) {
      after_colon_int(line, ts_in);
    } else if (line.find("# fields:") != std::string::npos) {
      break;
    }
  }
  if (ver != 1) return 3;
  if (n_in != ncells) return 4;
  static const int width[7] = {3, 1, 1, 3, 3, 1, 1};
  for (int k = 0; k < 7; k++) {
    // list-directed: values from successive lines until the block is full,
    // the rest of the last line discarded; then the '# end' line is skipped
    const size_t need = (size_t)width[k] * ncells;
    size_t got = 0;
    while (got < need) {
      if (!getline(line)) return 5;
      const char *p = line.c_str();
      while (*p && got < need) {
        while (*p == ' ' || *p == '\t' || *p == ',' || *p == '\r') p++;
        if (!*p) break;
        const char *q = p;
        while (*q && *q != ' ' && *q != '\t' && *q != ',' && *q != '\r') q++;
        double v;
        if (!parse_real(p, (size_t)(q - p), v)) return 5;
        f[k][got++] = v;
        p = q;
      }
    }
    getline(line);                           // '# end <name>'
  }
  *ts_order = ts_in;
  return 0;
}

}  // namespace ms
