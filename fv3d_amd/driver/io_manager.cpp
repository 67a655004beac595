// io_manager.cpp -- snapshot output with the reference's file layout
// (IOManager.h:88-191), written with the HDF5 C API.
//
// <out>.h5:
//   root int attributes Ntx Nty Ntz Nx Ny Nz ibeg iend jbeg jend kbeg kend,
//   root string attribute "problem" (variable-length string),
//   root datasets x, y, z: vertex coordinates (Nz+1)(Ny+1)(Nx+1), k-j-i
//     order, values (i-ibeg)*dx etc. (xmin is not added, as in the reference),
//   group ite_<n>: datasets rho u v w prs, (Nz, Ny, Nx) f64, attribute time.
// <out>.xdmf: a temporal collection rewritten at every snapshot by seeking
//   back over the footer.  The text is the reference's byte for byte
//   (including the seek over the footer's terminating NUL), so tools that
//   read the reference's XDMF read this one.
#include "io_manager.h"

#include <hdf5.h>

#include <cstdarg>
#include <cstdio>
#include <stdexcept>
#include <string>
#include <vector>

namespace fv3d {

namespace {

std::string fmt(const char* f, ...) __attribute__((format(printf, 1, 2)));
std::string fmt(const char* f, ...) {
  va_list ap, ap2;
  va_start(ap, f);
  va_copy(ap2, ap);
  const int n = std::vsnprintf(nullptr, 0, f, ap);
  va_end(ap);
  std::string s((size_t)n + 1, '\0');
  std::vsnprintf(&s[0], s.size(), f, ap2);
  va_end(ap2);
  s.resize((size_t)n);
  return s;
}

const char kFooter[] = "\n  </Grid>\n</Domain>\n</Xdmf>";

std::string xdmf_header(const fv3d_params& p, const std::string& h5) {
  const int z = p.Nz + 1, y = p.Ny + 1, x = p.Nx + 1;
  std::string s =
      "<?xml version=\"1.0\" ?>\n"
      "<!DOCTYPE Xdmf SYSTEM \"Xdmf.dtd\" []>\n"
      "<Xdmf Version=\"2.0\">\n"
      "<Domain CollectionType=\"Temporal\">\n"
      "  <Grid Name=\"MainTimeSeries\" GridType=\"Collection\" CollectionType=\"Temporal\">\n";
  s += fmt("    <Topology Name=\"Main Topology\" TopologyType=\"3DSMesh\" NumberOfElements=\"%d %d %d\"/>\n",
           z, y, x);
  s += "    <Geometry Name=\"Main Geometry\" GeometryType=\"X_Y_Z\">\n";
  for (const char* c : {"x", "y", "z"})
    s += fmt("      <DataItem Dimensions=\"%d %d %d\" NumberType=\"Float\" Precision=\"8\" "
             "Format=\"HDF\">%s:/%s</DataItem>\n",
             z, y, x, h5.c_str(), c);
  s += "    </Geometry>\n    ";
  return s;
}

std::string xdmf_field(const fv3d_params& p, const std::string& h5, const std::string& group,
                       const char* name) {
  return fmt("\n      <Attribute Name=\"%s\" AttributeType=\"Scalar\" Center=\"Cell\">"
             "\n        <DataItem Dimensions=\"%d %d %d\" NumberType=\"Float\" Precision=\"8\" "
             "Format=\"HDF\">%s:/%s/%s</DataItem>\n      </Attribute>",
             name, p.Nz, p.Ny, p.Nx, h5.c_str(), group.c_str(), name);
}

std::string xdmf_velocity(const fv3d_params& p, const std::string& h5, const std::string& group) {
  std::string s = fmt("\n      <Attribute Name=\"velocity\" AttributeType=\"Vector\" Center=\"Cell\">"
                      "\n        <DataItem Dimensions=\"%d %d %d 3\" ItemType=\"Function\" "
                      "Function=\"JOIN($0, $1, $2)\">",
                      p.Nz, p.Ny, p.Nx);
  for (const char* c : {"u", "v", "w"})
    s += fmt("\n          <DataItem Dimensions=\"%d %d %d\" NumberType=\"Float\" Precision=\"8\" "
             "Format=\"HDF\">%s:/%s/%s</DataItem>",
             p.Nz, p.Ny, p.Nx, h5.c_str(), group.c_str(), c);
  s += "\n        </DataItem>\n      </Attribute>";
  return s;
}

void h5check(herr_t e, const char* what) {
  if (e < 0) throw std::runtime_error(std::string("HDF5: ") + what + " failed");
}

void int_attr(hid_t loc, const char* name, int v) {
  hid_t sp = H5Screate(H5S_SCALAR);
  hid_t a = H5Acreate2(loc, name, H5T_NATIVE_INT, sp, H5P_DEFAULT, H5P_DEFAULT);
  h5check(H5Awrite(a, H5T_NATIVE_INT, &v), name);
  H5Aclose(a);
  H5Sclose(sp);
}

void double_attr(hid_t loc, const char* name, double v) {
  hid_t sp = H5Screate(H5S_SCALAR);
  hid_t a = H5Acreate2(loc, name, H5T_NATIVE_DOUBLE, sp, H5P_DEFAULT, H5P_DEFAULT);
  h5check(H5Awrite(a, H5T_NATIVE_DOUBLE, &v), name);
  H5Aclose(a);
  H5Sclose(sp);
}

void string_attr(hid_t loc, const char* name, const std::string& v) {
  hid_t t = H5Tcopy(H5T_C_S1);
  H5Tset_size(t, H5T_VARIABLE);
  hid_t sp = H5Screate(H5S_SCALAR);
  hid_t a = H5Acreate2(loc, name, t, sp, H5P_DEFAULT, H5P_DEFAULT);
  const char* c = v.c_str();
  h5check(H5Awrite(a, t, &c), name);
  H5Aclose(a);
  H5Sclose(sp);
  H5Tclose(t);
}

void dataset(hid_t loc, const char* name, int rank, const hsize_t* dims, const double* data) {
  hid_t sp = H5Screate_simple(rank, dims, nullptr);
  hid_t d = H5Dcreate2(loc, name, H5T_NATIVE_DOUBLE, sp, H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT);
  h5check(H5Dwrite(d, H5T_NATIVE_DOUBLE, H5S_ALL, H5S_ALL, H5P_DEFAULT, data), name);
  H5Dclose(d);
  H5Sclose(sp);
}

}  // namespace

IOManager::IOManager(const RunConfig& cfg) : cfg_(cfg) {}

void IOManager::saveSolution(const HostArray& Q, int iteration, double t, double /*dt*/) {
  const fv3d_params& p = cfg_.params;
  const std::string path = "ite_" + std::to_string(iteration);
  const std::string h5 = cfg_.filename_out + ".h5";
  const std::string xdmf = cfg_.filename_out + ".xdmf";

  hid_t file = iteration == 0 ? H5Fcreate(h5.c_str(), H5F_ACC_TRUNC, H5P_DEFAULT, H5P_DEFAULT)
                              : H5Fopen(h5.c_str(), H5F_ACC_RDWR, H5P_DEFAULT);
  if (file < 0) throw std::runtime_error("cannot open " + h5);
  FILE* fx = std::fopen(xdmf.c_str(), iteration == 0 ? "w+" : "r+");
  if (!fx) throw std::runtime_error("cannot open " + xdmf);

  if (iteration == 0) {
    int_attr(file, "Ntx", p.Ntx); int_attr(file, "Nty", p.Nty); int_attr(file, "Ntz", p.Ntz);
    int_attr(file, "Nx", p.Nx); int_attr(file, "Ny", p.Ny); int_attr(file, "Nz", p.Nz);
    int_attr(file, "ibeg", p.ibeg); int_attr(file, "iend", p.iend);
    int_attr(file, "jbeg", p.jbeg); int_attr(file, "jend", p.jend);
    int_attr(file, "kbeg", p.kbeg); int_attr(file, "kend", p.kend);
    string_attr(file, "problem", cfg_.problem);
    std::vector<double> x, y, z;
    for (int k = p.kbeg; k <= p.kend; ++k)
      for (int j = p.jbeg; j <= p.jend; ++j)
        for (int i = p.ibeg; i <= p.iend; ++i) {
          x.push_back((i - p.ibeg) * p.dx);
          y.push_back((j - p.jbeg) * p.dy);
          z.push_back((k - p.kbeg) * p.dz);
        }
    const hsize_t n = x.size();
    dataset(file, "x", 1, &n, x.data());
    dataset(file, "y", 1, &n, y.data());
    dataset(file, "z", 1, &n, z.data());
    const std::string head = xdmf_header(p, h5);
    std::fwrite(head.data(), 1, head.size(), fx);
    std::fwrite(kFooter, 1, sizeof(kFooter) - 1, fx);
  }

  // domain cells of each field, (Nz, Ny, Nx)
  const size_t nd = (size_t)p.Nx * p.Ny * p.Nz;
  std::vector<double> f[5];
  for (auto& v : f) v.resize(nd);
  size_t o = 0;
  for (int k = p.kbeg; k < p.kend; ++k)
    for (int j = p.jbeg; j < p.jend; ++j)
      for (int i = p.ibeg; i < p.iend; ++i, ++o) {
        const double* q = &Q[(((size_t)k * p.Nty + j) * p.Ntx + i) * 5];
        for (int v = 0; v < 5; ++v) f[v][o] = q[v];
      }
  hid_t g = H5Gcreate2(file, path.c_str(), H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT);
  if (g < 0) throw std::runtime_error("cannot create group " + path);
  const hsize_t dims[3] = {(hsize_t)p.Nz, (hsize_t)p.Ny, (hsize_t)p.Nx};
  dataset(g, "rho", 3, dims, f[FV3D_IR].data());
  dataset(g, "u", 3, dims, f[FV3D_IU].data());
  dataset(g, "v", 3, dims, f[FV3D_IV].data());
  dataset(g, "w", 3, dims, f[FV3D_IW].data());
  dataset(g, "prs", 3, dims, f[FV3D_IP].data());
  double_attr(g, "time", t);
  H5Gclose(g);
  H5Fclose(file);

  // the reference seeks back sizeof(footer), which counts the terminating NUL
  std::fseek(fx, -(long)sizeof(kFooter), SEEK_END);
  std::string s = fmt("\n    <Grid Name=\"Cells\" GridType=\"Uniform\">"
                      "\n      <Time TimeType=\"Single\" Value=\"%lf\" />"
                      "\n      <Topology Reference=\"//Topology[@Name='Main Topology']\" />"
                      "\n      <Geometry Reference=\"//Geometry[@Name='Main Geometry']\" />",
                      t);
  s += xdmf_field(p, h5, path, "rho");
  s += xdmf_velocity(p, h5, path);
  s += xdmf_field(p, h5, path, "prs");
  s += "\n    </Grid>\n  ";
  std::fwrite(s.data(), 1, s.size(), fx);
  std::fwrite(kFooter, 1, sizeof(kFooter) - 1, fx);
  std::fclose(fx);
}

}  // namespace fv3d
