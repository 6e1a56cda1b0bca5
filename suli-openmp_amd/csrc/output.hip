// Field output of the reference's plot_f and its ParaView index.
//
// plot_f (4_plot.f90:260-363) writes, every frame, one Fortran unformatted
// sequential record arquivos/campos_NNNNN holding node-averaged velocities,
// the level set, the pressure, the vorticity and the face viscosities in
// FP32; visu (12_paraview_3D.f90:1-179) writes arquivos/visu.xdmf, which
// points ParaView at the arrays inside those records.  Here one kernel
// (k_frame_record) builds the record on the device from the resident fields
// and one copy brings it to the host; the file is the reference's byte for
// byte (4-byte record length before and after the payload, little endian).
#include <cstdio>
#include <string>
#include <vector>

#include "ctx.hpp"
#include "launch.hpp"

#define NEEDCTX()                                              \
  do {                                                         \
    if (!g_ctx) {                                              \
      suli_set_error("library not initialised (suli_init)");   \
      return 2;                                                \
    }                                                          \
  } while (0)

namespace {

int64_t frame_values(int64_t nx, int64_t ny, int64_t nz, int t_plot) {
  const int64_t n1 = (nx + 1) * (ny + 1) * (nz + 1), n = nx * ny * nz;
  int64_t v = 6 * n1 + 2 * n;
  if (t_plot == 1) v += 4 * n + (nx + 1) * ny * nz + nx * (ny + 1) * nz + nx * ny * (nz + 1);
  return v;
}

bool t_plot_ok(int t_plot, const char *fn) {
  if (t_plot == 0 || t_plot == 1) return true;
  suli_set_error(std::string(fn) + ": t_plot must be 0 (velocity, level set) or 1 (adds pressure, vorticity, "
                 "face viscosities)");
  return false;
}

// Coordinates of the record: x1(0) = -dx and x1(i) = x1(i-1) + dx as a
// running FP64 sum (4_plot.f90:264-266,275-277), i = 1..n+1, rounded to FP32.
void axis_table(int n, double h, float *t) {
  double x = -h;
  t[0] = (float)x;
  for (int i = 1; i <= n + 1; i++) {
    x = x + h;
    t[i] = (float)x;
  }
}

// The record on the device (c.out_dev), n values.
int frame_build(Ctx &c, int t_plot, int64_t *nval) {
  const Grid &g = c.g;
  const int64_t n = frame_values(g.nx, g.ny, g.nz, t_plot);
  *nval = n;
  if (n > c.out_cap) {
    if (c.out_dev) (void)hipFree(c.out_dev);
    c.out_dev = nullptr;
    c.out_cap = 0;
    HIPCHK(hipMalloc((void **)&c.out_dev, (size_t)n * sizeof(float)));
    c.out_cap = n;
  }
  const int nt = (g.nx + 2) + (g.ny + 2) + (g.nz + 2);
  if (!c.out_tab) {
    HIPCHK(hipMalloc((void **)&c.out_tab, (size_t)nt * sizeof(float)));
    c.allocs.push_back(c.out_tab);
    std::vector<float> t((size_t)nt);
    axis_table(g.nx, g.dx, t.data());
    axis_table(g.ny, g.dy, t.data() + g.nx + 2);
    axis_table(g.nz, g.dz, t.data() + g.nx + 2 + g.ny + 2);
    HIPCHK(hipMemcpyAsync(c.out_tab, t.data(), (size_t)nt * sizeof(float), hipMemcpyHostToDevice, c.st));
    HIPCHK(hipStreamSynchronize(c.st));
  }
  const float *xs = c.out_tab, *ys = xs + g.nx + 2, *zs = ys + g.ny + 2;
  DISPATCH(c, Launch<TT>::frame_record(g, c.st, t_plot, c.u, c.v, c.w, c.ls, c.prd1, c.nut, c.obst ? c.kw_d : nullptr, xs,
                                               ys, zs, c.out_dev));
  HIPCHK(hipGetLastError());
  return 0;
}

}  // namespace

extern "C" int suli_frame_values(int t_plot, int64_t *n) {
  NEEDCTX();
  if (!t_plot_ok(t_plot, "suli_frame_values")) return 1;
  *n = frame_values(g_ctx->g.nx, g_ctx->g.ny, g_ctx->g.nz, t_plot);
  return 0;
}

extern "C" int suli_frame_record(int t_plot, float *buf, int64_t n) {
  NEEDCTX();
  Ctx &c = *g_ctx;
  if (!t_plot_ok(t_plot, "suli_frame_record")) return 1;
  const int64_t need = frame_values(c.g.nx, c.g.ny, c.g.nz, t_plot);
  if (n != need) {
    suli_set_error("suli_frame_record: the record holds " + std::to_string(need) + " values, got " +
                   std::to_string(n));
    return 1;
  }
  int64_t nv = 0;
  if (frame_build(c, t_plot, &nv)) return 1;
  HIPCHK(hipMemcpyAsync(buf, c.out_dev, (size_t)nv * sizeof(float), hipMemcpyDeviceToHost, c.st));
  HIPCHK(hipStreamSynchronize(c.st));
  return 0;
}

extern "C" int suli_write_frame(const char *path, int t_plot) {
  NEEDCTX();
  Ctx &c = *g_ctx;
  if (!t_plot_ok(t_plot, "suli_write_frame")) return 1;
  if (!path) {
    suli_set_error("suli_write_frame: null path");
    return 1;
  }
  const int64_t nv = frame_values(c.g.nx, c.g.ny, c.g.nz, t_plot);
  const int64_t bytes = nv * (int64_t)sizeof(float);
  if (bytes > 2147483647LL) {  // one record with a 4-byte length
    suli_set_error("suli_write_frame: the record (" + std::to_string(bytes) +
                   " bytes) exceeds the 2^31-1 bytes of a 4-byte record length");
    return 1;
  }
  int64_t n = 0;
  if (frame_build(c, t_plot, &n)) return 1;
  if (n > c.out_host_cap) {
    if (c.out_host) (void)hipHostFree(c.out_host);
    c.out_host = nullptr;
    c.out_host_cap = 0;
    HIPCHK(hipHostMalloc((void **)&c.out_host, (size_t)n * sizeof(float), hipHostMallocDefault));
    c.out_host_cap = n;
  }
  HIPCHK(hipMemcpyAsync(c.out_host, c.out_dev, (size_t)bytes, hipMemcpyDeviceToHost, c.st));
  HIPCHK(hipStreamSynchronize(c.st));
  FILE *f = std::fopen(path, "wb");
  if (!f) {
    suli_set_error(std::string("suli_write_frame: cannot open ") + path);
    return 1;
  }
  const int32_t len = (int32_t)bytes;
  const bool ok = std::fwrite(&len, 4, 1, f) == 1 && std::fwrite(c.out_host, 1, (size_t)bytes, f) == (size_t)bytes &&
                  std::fwrite(&len, 4, 1, f) == 1;
  if (std::fclose(f) != 0 || !ok) {
    suli_set_error(std::string("suli_write_frame: write failed on ") + path);
    return 1;
  }
  return 0;
}

// visu (12_paraview_3D.f90:1-179): the XDMF index of the records
// campos_<first>..campos_<last> (5 digits, in the index's directory), one
// uniform 3DSMesh grid per record inside a temporal collection starting at
// 0 with stride dt_frame.  Names, dimensions and byte offsets (Seek) are
// the reference's, its two quirks included: the time hyperslab lists start
// and stride only, and the face-viscosity components are addressed as if
// each held nx*ny*nz values.  Host-only: needs no context.
extern "C" int suli_write_xdmf(const char *path, int nx, int ny, int nz, int first, int last, int t_plot,
                               double dt_frame) {
  if (!t_plot_ok(t_plot, "suli_write_xdmf")) return 1;
  if (!path || nx < 1 || ny < 1 || nz < 1 || first < 0 || last < first || last > 99999) {
    suli_set_error("suli_write_xdmf: needs a path, nx, ny, nz >= 1 and 0 <= first <= last <= 99999");
    return 1;
  }
  FILE *f = std::fopen(path, "w");
  if (!f) {
    suli_set_error(std::string("suli_write_xdmf: cannot open ") + path);
    return 1;
  }
  const long long n1 = (long long)(nx + 1) * (ny + 1) * (nz + 1), n = (long long)nx * ny * nz;
  const int nx1 = nx + 1, ny1 = ny + 1, nz1 = nz + 1;
  std::fprintf(f, "<?xml version=\"1.0\" ?>\n<!DOCTYPE Xdmf SYSTEM \"Xdmf.dtd\" []>\n");
  std::fprintf(f, "<Xdmf xmlns:xi=\"http://www.w3.org/2001/XInclude\" Version=\"2.0\">\n<Domain>\n");
  std::fprintf(f, "  <Grid Name=\"TimeSeries\" GridType=\"Collection\" CollectionType=\"Temporal\">\n");
  std::fprintf(f, "    <Time TimeType=\"HyperSlab\">\n");
  std::fprintf(f, "      <DataItem Format=\"XML\" NumberType=\"Float\" Dimensions=\"3\">\n");
  std::fprintf(f, "        <!--Start, Stride, Count-->\n        0.0 %.17g\n      </DataItem>\n    </Time>\n", dt_frame);
  auto item = [&](const char *indent, const char *name, long long seek, const char *dims, const char *file) {
    std::fprintf(f, "%s<DataItem%s%s%s Format=\"Binary\" DataType=\"Float\" Precision=\"4\" Endian=\"little\" "
                    "Seek=\"%lld\" Dimensions=\"%s\">\n%s  %s\n%s</DataItem>\n",
                 indent, name ? " Name=\"" : "", name ? name : "", name ? "\"" : "", seek, dims, indent, file, indent);
  };
  char nodes[64], nodes1[64], cells[64], cells1[64];
  std::snprintf(nodes, sizeof nodes, "%d %d %d", nz1, ny1, nx1);
  std::snprintf(nodes1, sizeof nodes1, "%d %d %d 1", nz1, ny1, nx1);
  std::snprintf(cells, sizeof cells, "%d %d %d", nz, ny, nx);
  std::snprintf(cells1, sizeof cells1, "%d %d %d 1", nz, ny, nx);
  for (int q = first; q <= last; q++) {
    char file[32];
    std::snprintf(file, sizeof file, "campos_%05d", q);
    std::fprintf(f, "    <Grid Name=\"%s\" GridType=\"Uniform\">\n", file);
    std::fprintf(f, "      <Topology name=\"topo\" TopologyType=\"3DSMesh\" Dimensions=\"%s\">\n      </Topology>\n",
                 nodes);
    std::fprintf(f, "      <Geometry name=\"geo\" Type=\"X_Y_Z\">\n");
    item("        ", "X", n1 * 4 * 0 + 4, nodes, file);
    item("        ", "Y", n1 * 4 * 1 + 4, nodes, file);
    item("        ", "Z", n1 * 4 * 2 + 4, nodes, file);
    std::fprintf(f, "      </Geometry>\n");
    auto vector = [&](const char *name, const char *center, const char *dims3, const char *dims1, long long s0,
                      long long step) {
      std::fprintf(f, "      <Attribute Name=\"%s\" AttributeType=\"Vector\" Center=\"%s\">\n", name, center);
      std::fprintf(f, "        <DataItem Dimensions=\"%s 3\" Function=\"JOIN($0, $1, $2)\" ItemType=\"Function\">\n",
                   dims3);
      for (int c = 0; c < 3; c++) item("          ", nullptr, s0 + c * step, dims1, file);
      std::fprintf(f, "        </DataItem>\n      </Attribute>\n");
    };
    auto scalar = [&](const char *name, long long seek) {
      std::fprintf(f, "      <Attribute Name=\"%s\" Center=\"Cell\">\n", name);
      item("        ", nullptr, seek, cells, file);
      std::fprintf(f, "      </Attribute>\n");
    };
    vector("V", "Node", nodes, nodes1, n1 * 4 * 3 + 4, n1 * 4);
    scalar("level_set", n1 * 4 * 6 + 4);
    scalar("obstaculo", n1 * 4 * 6 + n * 4 * 1 + 4);
    if (t_plot == 1) {
      scalar("pressure", n1 * 4 * 6 + n * 4 * 2 + 4);
      vector("vorticity", "Cell", cells, cells1, n1 * 4 * 6 + n * 4 * 3 + 4, n * 4);
      vector("visc", "Cell", cells, cells1, n1 * 4 * 6 + n * 4 * 6 + 4, n * 4);
    }
    std::fprintf(f, "    </Grid>\n");
  }
  std::fprintf(f, "  </Grid>\n</Domain>\n</Xdmf>\n");
  if (std::fclose(f) != 0) {
    suli_set_error(std::string("suli_write_xdmf: write failed on ") + path);
    return 1;
  }
  return 0;
}
