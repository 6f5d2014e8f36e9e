// HDF5 + XDMF output and restart of the conservative state (libg5h5.so).  Host code over the
// HDF5 1.10 C library; the state is the host copy U[block][var][cell] of the C ABI.
//
// Write: per variable, the blocks are scattered into one buffer in global cell order (OpenMP over
// blocks), which is written as one contiguous dataset; read is the inverse gather.
#include <hdf5.h>

#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/godunov_five_eq_io.h"

namespace
{

thread_local std::string g_err;

int
fail(int rc, std::string const & msg)
{
  g_err = msg;
  return rc;
}

// models/FiveEq.h:93-135: hydro names, then alpha_<mat>, alpha_rho_<mat>
const char * const kNames2[8] = { "rho_mix", "e_tot", "rho_vx", "rho_vy", "alpha_0", "alpha_rho_0", "alpha_1",
                                  "alpha_rho_1" };
const char * const kNames3[9] = { "rho_mix", "e_tot",   "rho_vx",  "rho_vy",     "rho_vz",
                                  "alpha_0", "alpha_rho_0", "alpha_1", "alpha_rho_1" };

struct Mesh
{
  int     dim = 2, b = 16, nvar = 8;
  int     nb[3] = { 1, 1, 1 };
  int64_t ncell_blk = 256, nblocks = 0, ncells = 0;
  int64_t n[3] = { 1, 1, 1 }; // cells per direction
};

int
mesh_of(g5_config const * cfg, Mesh & m)
{
  if (cfg == nullptr)
    return fail(G5_ERR_INVALID, "null configuration");
  if (cfg->dim != 2 and cfg->dim != 3)
    return fail(G5_ERR_INVALID, "dim must be 2 or 3");
  if (cfg->bsize < 1)
    return fail(G5_ERR_INVALID, "block size must be positive");
  m.dim = cfg->dim;
  m.b = cfg->bsize;
  m.nvar = m.dim == 2 ? 8 : 9;
  m.ncell_blk = m.dim == 2 ? int64_t(m.b) * m.b : int64_t(m.b) * m.b * m.b;
  m.nblocks = 1;
  m.ncells = 1;
  for (int d = 0; d < 3; ++d)
  {
    m.nb[d] = d < m.dim ? cfg->nb[d] : 1;
    if (m.nb[d] < 1)
      return fail(G5_ERR_INVALID, "block counts must be positive");
    m.n[d] = d < m.dim ? int64_t(m.nb[d]) * m.b : 1;
    m.nblocks *= m.nb[d];
    m.ncells *= m.n[d];
  }
  return G5_OK;
}

// global cell index of cell c of block k (brick order, x fastest; cells x fastest)
inline int64_t
global_index(Mesh const & m, int64_t k, int64_t c)
{
  const int64_t bx = k % m.nb[0], by = (k / m.nb[0]) % m.nb[1], bz = k / (int64_t(m.nb[0]) * m.nb[1]);
  const int64_t cx = c % m.b, cy = (c / m.b) % m.b, cz = c / (int64_t(m.b) * m.b);
  return ((bz * m.b + cz) * m.n[1] + (by * m.b + cy)) * m.n[0] + (bx * m.b + cx);
}

void
scatter_var(Mesh const & m, double const * U, int v, double * G)
{
#pragma omp parallel for schedule(static)
  for (int64_t k = 0; k < m.nblocks; ++k)
  {
    double const * src = U + (k * m.nvar + v) * m.ncell_blk;
    // rows of b cells along x are contiguous on both sides
    for (int64_t c = 0; c < m.ncell_blk; c += m.b)
      std::memcpy(G + global_index(m, k, c), src + c, sizeof(double) * m.b);
  }
}

void
gather_var(Mesh const & m, double const * G, int v, double * U)
{
#pragma omp parallel for schedule(static)
  for (int64_t k = 0; k < m.nblocks; ++k)
  {
    double * dst = U + (k * m.nvar + v) * m.ncell_blk;
    for (int64_t c = 0; c < m.ncell_blk; c += m.b)
      std::memcpy(dst + c, G + global_index(m, k, c), sizeof(double) * m.b);
  }
}

// RAII over an hid_t with its close function
struct H
{
  hid_t id;
  herr_t (*close)(hid_t);
  H(hid_t i, herr_t (*c)(hid_t)) : id(i), close(c) {}
  ~H()
  {
    if (id >= 0)
      close(id);
  }
  H(H const &) = delete;
  H & operator=(H const &) = delete;
  bool
  bad() const
  {
    return id < 0;
  }
};

bool
write_attr(hid_t loc, char const * name, hid_t file_type, hid_t mem_type, int rank, hsize_t len, void const * data)
{
  H sp(rank == 0 ? H5Screate(H5S_SCALAR) : H5Screate_simple(1, &len, nullptr), H5Sclose);
  if (sp.bad())
    return false;
  H a(H5Acreate2(loc, name, file_type, sp.id, H5P_DEFAULT, H5P_DEFAULT), H5Aclose);
  return !a.bad() and H5Awrite(a.id, mem_type, data) >= 0;
}

bool
read_attr(hid_t loc, char const * name, hid_t mem_type, hsize_t len, void * data)
{
  if (H5Aexists(loc, name) <= 0)
    return false;
  H a(H5Aopen(loc, name, H5P_DEFAULT), H5Aclose);
  if (a.bad())
    return false;
  H sp(H5Aget_space(a.id), H5Sclose);
  const hssize_t npts = H5Sget_simple_extent_npoints(sp.id);
  if (npts != static_cast<hssize_t>(len))
    return false;
  return H5Aread(a.id, mem_type, data) >= 0;
}

std::string
base_name(std::string const & path)
{
  const auto k = path.find_last_of('/');
  return k == std::string::npos ? path : path.substr(k + 1);
}

int
write_xdmf(std::string const & path, std::string const & h5file, Mesh const & m, g5_config const * cfg, double t,
           std::vector<int> const & vars)
{
  FILE * f = std::fopen(path.c_str(), "w");
  if (!f)
    return fail(G5_ERR_IO, "cannot create " + path);
  // CoRectMesh: point dimensions, origin and spacing listed slowest axis first
  std::fprintf(f, "<?xml version=\"1.0\" ?>\n<!DOCTYPE Xdmf SYSTEM \"Xdmf.dtd\" []>\n<Xdmf Version=\"2.0\">\n");
  std::fprintf(f, " <Domain>\n  <Grid Name=\"mesh\" GridType=\"Uniform\">\n");
  std::fprintf(f, "   <Time Value=\"%.17g\"/>\n", t);
  if (m.dim == 2)
  {
    std::fprintf(f, "   <Topology TopologyType=\"2DCoRectMesh\" Dimensions=\"%lld %lld\"/>\n",
                 static_cast<long long>(m.n[1] + 1), static_cast<long long>(m.n[0] + 1));
    std::fprintf(f, "   <Geometry GeometryType=\"ORIGIN_DXDY\">\n");
    std::fprintf(f, "    <DataItem Dimensions=\"2\" NumberType=\"Float\" Precision=\"8\" Format=\"XML\">%.17g %.17g</DataItem>\n",
                 cfg->xmin[1], cfg->xmin[0]);
    std::fprintf(f, "    <DataItem Dimensions=\"2\" NumberType=\"Float\" Precision=\"8\" Format=\"XML\">%.17g %.17g</DataItem>\n",
                 cfg->dx, cfg->dx);
  }
  else
  {
    std::fprintf(f, "   <Topology TopologyType=\"3DCoRectMesh\" Dimensions=\"%lld %lld %lld\"/>\n",
                 static_cast<long long>(m.n[2] + 1), static_cast<long long>(m.n[1] + 1),
                 static_cast<long long>(m.n[0] + 1));
    std::fprintf(f, "   <Geometry GeometryType=\"ORIGIN_DXDYDZ\">\n");
    std::fprintf(f,
                 "    <DataItem Dimensions=\"3\" NumberType=\"Float\" Precision=\"8\" Format=\"XML\">%.17g %.17g %.17g</DataItem>\n",
                 cfg->xmin[2], cfg->xmin[1], cfg->xmin[0]);
    std::fprintf(f,
                 "    <DataItem Dimensions=\"3\" NumberType=\"Float\" Precision=\"8\" Format=\"XML\">%.17g %.17g %.17g</DataItem>\n",
                 cfg->dx, cfg->dx, cfg->dx);
  }
  std::fprintf(f, "   </Geometry>\n");
  const std::string dims = m.dim == 2 ? std::to_string(m.n[1]) + " " + std::to_string(m.n[0])
                                      : std::to_string(m.n[2]) + " " + std::to_string(m.n[1]) + " " +
                                          std::to_string(m.n[0]);
  for (int v : vars)
  {
    const char * name = g5_h5_var_name(m.dim, v);
    std::fprintf(f, "   <Attribute Name=\"%s\" AttributeType=\"Scalar\" Center=\"Cell\">\n", name);
    std::fprintf(f, "    <DataItem Dimensions=\"%s\" NumberType=\"Float\" Precision=\"8\" Format=\"HDF\">%s:/%s</DataItem>\n",
                 dims.c_str(), h5file.c_str(), name);
    std::fprintf(f, "   </Attribute>\n");
  }
  std::fprintf(f, "  </Grid>\n </Domain>\n</Xdmf>\n");
  return std::fclose(f) == 0 ? G5_OK : fail(G5_ERR_IO, "cannot write " + path);
}

} // namespace

extern "C" const char *
g5_h5_var_name(int dim, int v)
{
  if (dim == 2 and v >= 0 and v < 8)
    return kNames2[v];
  if (dim == 3 and v >= 0 and v < 9)
    return kNames3[v];
  return nullptr;
}

extern "C" const char *
g5_h5_last_error(void)
{
  return g_err.c_str();
}

extern "C" int
g5_h5_write(const char * dir, const char * basename, const g5_config * cfg, const double * U, size_t n, double t,
            int64_t iteration, uint32_t var_mask)
{
  Mesh m;
  int  rc = mesh_of(cfg, m);
  if (rc != G5_OK)
    return rc;
  if (dir == nullptr or basename == nullptr or basename[0] == '\0' or U == nullptr)
    return fail(G5_ERR_INVALID, "null argument");
  if (n != static_cast<size_t>(m.nblocks * m.nvar * m.ncell_blk))
    return fail(G5_ERR_INVALID, "state size does not match the configuration");
  if (var_mask >> m.nvar)
    return fail(G5_ERR_INVALID, "variable mask selects a variable the state does not have");
  std::vector<int> vars;
  for (int v = 0; v < m.nvar; ++v)
    if (var_mask == 0 or (var_mask >> v) & 1u)
      vars.push_back(v);

  const std::string stem = std::string(dir) + (dir[0] ? "/" : "") + basename;
  const std::string h5path = stem + ".h5";
  H5Eset_auto2(H5E_DEFAULT, nullptr, nullptr);
  H file(H5Fcreate(h5path.c_str(), H5F_ACC_TRUNC, H5P_DEFAULT, H5P_DEFAULT), H5Fclose);
  if (file.bad())
    return fail(G5_ERR_IO, "cannot create " + h5path);
  {
    H run(H5Gcreate2(file.id, "run", H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT), H5Gclose);
    if (run.bad() or !write_attr(run.id, "time", H5T_IEEE_F64LE, H5T_NATIVE_DOUBLE, 0, 1, &t) or
        !write_attr(run.id, "iteration", H5T_STD_I64LE, H5T_NATIVE_INT64, 0, 1, &iteration))
      return fail(G5_ERR_IO, "cannot write the run attributes of " + h5path);
    H           mesh(H5Gcreate2(file.id, "mesh", H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT), H5Gclose);
    const int32_t dim = m.dim, b = m.b;
    const int32_t nb[3] = { m.nb[0], m.nb[1], m.nb[2] };
    const double  xmin[3] = { cfg->xmin[0], cfg->xmin[1], cfg->xmin[2] };
    if (mesh.bad() or !write_attr(mesh.id, "dim", H5T_STD_I32LE, H5T_NATIVE_INT32, 0, 1, &dim) or
        !write_attr(mesh.id, "block_size", H5T_STD_I32LE, H5T_NATIVE_INT32, 0, 1, &b) or
        !write_attr(mesh.id, "nb", H5T_STD_I32LE, H5T_NATIVE_INT32, 1, 3, nb) or
        !write_attr(mesh.id, "dx", H5T_IEEE_F64LE, H5T_NATIVE_DOUBLE, 0, 1, &cfg->dx) or
        !write_attr(mesh.id, "xmin", H5T_IEEE_F64LE, H5T_NATIVE_DOUBLE, 1, 3, xmin))
      return fail(G5_ERR_IO, "cannot write the mesh attributes of " + h5path);
  }
  std::vector<double> G;
  try
  {
    G.resize(static_cast<size_t>(m.ncells));
  }
  catch (std::bad_alloc const &)
  {
    return fail(G5_ERR_NOMEM, "no host memory for one variable");
  }
  const hsize_t dims[3] = { static_cast<hsize_t>(m.dim == 2 ? m.n[1] : m.n[2]), static_cast<hsize_t>(m.dim == 2 ? m.n[0] : m.n[1]),
                            static_cast<hsize_t>(m.n[0]) };
  H space(H5Screate_simple(m.dim, dims, nullptr), H5Sclose);
  if (space.bad())
    return fail(G5_ERR_IO, "cannot create a dataspace");
  for (int v : vars)
  {
    scatter_var(m, U, v, G.data());
    const char * name = g5_h5_var_name(m.dim, v);
    H            ds(H5Dcreate2(file.id, name, H5T_IEEE_F64LE, space.id, H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT), H5Dclose);
    if (ds.bad() or H5Dwrite(ds.id, H5T_NATIVE_DOUBLE, H5S_ALL, H5S_ALL, H5P_DEFAULT, G.data()) < 0)
      return fail(G5_ERR_IO, std::string("cannot write dataset ") + name + " of " + h5path);
  }
  if (H5Fflush(file.id, H5F_SCOPE_LOCAL) < 0)
    return fail(G5_ERR_IO, "cannot flush " + h5path);
  return write_xdmf(stem + ".xmf", base_name(h5path), m, cfg, t, vars);
}

extern "C" int
g5_h5_read(const char * path, const g5_config * cfg, double * U, size_t n, double * t, int64_t * iteration)
{
  Mesh m;
  int  rc = mesh_of(cfg, m);
  if (rc != G5_OK)
    return rc;
  if (path == nullptr or U == nullptr or t == nullptr or iteration == nullptr)
    return fail(G5_ERR_INVALID, "null argument");
  if (n != static_cast<size_t>(m.nblocks * m.nvar * m.ncell_blk))
    return fail(G5_ERR_INVALID, "state size does not match the configuration");
  H5Eset_auto2(H5E_DEFAULT, nullptr, nullptr);
  const std::string p(path);
  H                 file(H5Fopen(path, H5F_ACC_RDONLY, H5P_DEFAULT), H5Fclose);
  if (file.bad())
    return fail(G5_ERR_IO, "cannot open " + p);
  double  tt = 0.0;
  int64_t it = 0;
  {
    if (H5Lexists(file.id, "run", H5P_DEFAULT) <= 0)
      return fail(G5_ERR_IO, p + ": no run group");
    H run(H5Gopen2(file.id, "run", H5P_DEFAULT), H5Gclose);
    if (run.bad() or !read_attr(run.id, "time", H5T_NATIVE_DOUBLE, 1, &tt) or
        !read_attr(run.id, "iteration", H5T_NATIVE_INT64, 1, &it))
      return fail(G5_ERR_IO, p + ": cannot read the run attributes time and iteration");
    if (H5Lexists(file.id, "mesh", H5P_DEFAULT) <= 0)
      return fail(G5_ERR_IO, p + ": no mesh group");
    H       mesh(H5Gopen2(file.id, "mesh", H5P_DEFAULT), H5Gclose);
    int32_t dim = 0, b = 0, nb[3] = { 0, 0, 0 };
    double  dx = 0.0;
    if (mesh.bad() or !read_attr(mesh.id, "dim", H5T_NATIVE_INT32, 1, &dim) or
        !read_attr(mesh.id, "block_size", H5T_NATIVE_INT32, 1, &b) or
        !read_attr(mesh.id, "nb", H5T_NATIVE_INT32, 3, nb) or !read_attr(mesh.id, "dx", H5T_NATIVE_DOUBLE, 1, &dx))
      return fail(G5_ERR_IO, p + ": cannot read the mesh attributes");
    if (dim != m.dim or b != m.b or nb[0] != m.nb[0] or nb[1] != m.nb[1] or nb[2] != m.nb[2] or dx != cfg->dx)
      return fail(G5_ERR_INVALID, p + ": the file's mesh (dim, block size, blocks, dx) is not the configuration's");
  }
  std::vector<double> G;
  try
  {
    G.resize(static_cast<size_t>(m.ncells));
  }
  catch (std::bad_alloc const &)
  {
    return fail(G5_ERR_NOMEM, "no host memory for one variable");
  }
  for (int v = 0; v < m.nvar; ++v)
  {
    const char * name = g5_h5_var_name(m.dim, v);
    if (H5Lexists(file.id, name, H5P_DEFAULT) <= 0)
      return fail(G5_ERR_IO, p + ": no dataset " + name + " (a checkpoint holds every variable)");
    H ds(H5Dopen2(file.id, name, H5P_DEFAULT), H5Dclose);
    if (ds.bad())
      return fail(G5_ERR_IO, p + ": cannot open dataset " + name);
    H             sp(H5Dget_space(ds.id), H5Sclose);
    hsize_t       dims[3] = { 0, 0, 0 };
    const int     rank = H5Sget_simple_extent_ndims(sp.id);
    if (rank != m.dim or H5Sget_simple_extent_dims(sp.id, dims, nullptr) != m.dim)
      return fail(G5_ERR_INVALID, p + ": dataset " + name + " has the wrong rank");
    const hsize_t want[3] = { static_cast<hsize_t>(m.dim == 2 ? m.n[1] : m.n[2]),
                              static_cast<hsize_t>(m.dim == 2 ? m.n[0] : m.n[1]), static_cast<hsize_t>(m.n[0]) };
    for (int d = 0; d < m.dim; ++d)
      if (dims[d] != want[d])
        return fail(G5_ERR_INVALID, p + ": dataset " + name + " has the wrong shape");
    if (H5Dread(ds.id, H5T_NATIVE_DOUBLE, H5S_ALL, H5S_ALL, H5P_DEFAULT, G.data()) < 0)
      return fail(G5_ERR_IO, p + ": cannot read dataset " + name);
    gather_var(m, G.data(), v, U);
  }
  *t = tt;
  *iteration = it;
  return G5_OK;
}
