// Unstructured-grid output of patch data: VTU XML with appended raw binary,
// the content of PeanoClaw's VTUSubgridPlotter
// (grid/plotter/VTUSubgridPlotter.cpp:9-241), and legacy ASCII VTK, that of
// its VTKSubgridPlotter (VTKSubgridPlotter.cpp:9-241): one quadrilateral per subcell,
// vertices shared between subcells and patches, cell arrays q0..q2 (uNew),
// pWG0 (the parameter with ghost layer, b) and the metadata SubdivisionFactor,
// GhostlayerWidth, timeOld, timeNew, demandedMeshWidth and age.  Included at
// the end of pc_abi.hip (host code only).
#pragma once

#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <unordered_map>
#include <vector>

namespace {

// Vertex identity: coordinates rounded to 2^-40 of the unit (the positions of
// shared vertices computed from different patch origins agree far below it).
struct VtuKey {
  long long x, y;
  bool operator==(const VtuKey& o) const { return x == o.x && y == o.y; }
};
struct VtuKeyHash {
  size_t operator()(const VtuKey& k) const {
    return std::hash<long long>()(k.x) * 1000003u ^ std::hash<long long>()(k.y);
  }
};

constexpr int kVtuArrays = 10;
const char* const kVtuNames[kVtuArrays] = {"q0", "q1", "q2", "pWG0", "SubdivisionFactor",
                                           "GhostlayerWidth", "timeOld", "timeNew",
                                           "demandedMeshWidth", "age"};

// Vertices (x, y, z), quadrilaterals and cell arrays of the patches, as the
// reference's subgrid plotters lay them out (plotSubgrid / plotSubcell):
// with gap g a patch of level l is shrunk by 1 + g l around its centre
// (subcell size / (1 + g l), origin position + size * g l / 2); the VTU
// plotter uses g = 0, the VTK plotter g = 0.015 (VTKSubgridPlotter.cpp:22).
struct CellGrid {
  std::vector<double> pts;
  std::vector<int32_t> conn;
  std::vector<double> cell[kVtuArrays];
};

int collect_cells(int n, const pc_vtu_patch* pt, double gap, CellGrid& g, const char* who) {
  if (n < 0 || (n > 0 && !pt)) return fail(PC_E_INVALID, std::string(who) + ": invalid argument");
  std::unordered_map<VtuKey, int32_t, VtuKeyHash> index;
  const double scale = 1099511627776.0;  // 2^40
  for (int k = 0; k < n; ++k) {
    const pc_vtu_patch& P = pt[k];
    const int nx = P.subdivision_factor[0], ny = P.subdivision_factor[1];
    if (nx < 1 || ny < 1 || !P.q[0] || !P.q[1] || !P.q[2] || !P.b)
      return fail(PC_E_INVALID, std::string(who) + ": patch without data or with an empty grid");
    const double localGap = gap * P.level;
    const double sx = P.size[0] / nx / (1.0 + localGap), sy = P.size[1] / ny / (1.0 + localGap);
    const double x0 = P.position[0] + P.size[0] * localGap / 2.0;
    const double y0 = P.position[1] + P.size[1] * localGap / 2.0;
    std::vector<int32_t> vid((size_t)(nx + 1) * (ny + 1));
    for (int i = 0; i <= nx; ++i)
      for (int j = 0; j <= ny; ++j) {
        const double x = x0 + sx * i, y = y0 + sy * j;
        const VtuKey key{std::llround(x * scale), std::llround(y * scale)};
        auto it = index.find(key);
        int32_t id;
        if (it == index.end()) {
          id = (int32_t)(g.pts.size() / 3);
          index.emplace(key, id);
          g.pts.push_back(x);
          g.pts.push_back(y);
          g.pts.push_back(0.0);
        } else {
          id = it->second;
        }
        vid[(size_t)i * (ny + 1) + j] = id;
      }
    // subcells with x slowest and y fastest (the layout of the arrays)
    for (int i = 0; i < nx; ++i)
      for (int j = 0; j < ny; ++j) {
        const size_t v00 = (size_t)i * (ny + 1) + j;
        g.conn.push_back(vid[v00]);
        g.conn.push_back(vid[v00 + (ny + 1)]);
        g.conn.push_back(vid[v00 + (ny + 1) + 1]);
        g.conn.push_back(vid[v00 + 1]);
        const size_t c = (size_t)i * ny + j;
        g.cell[0].push_back(P.q[0][c]);
        g.cell[1].push_back(P.q[1][c]);
        g.cell[2].push_back(P.q[2][c]);
        g.cell[3].push_back(P.b[c]);
        g.cell[4].push_back(nx);
        g.cell[5].push_back(P.ghostlayer_width);
        g.cell[6].push_back(P.time);
        g.cell[7].push_back(P.time + P.timestep_size);
        g.cell[8].push_back(P.demanded_mesh_width);
        g.cell[9].push_back(P.age);
      }
  }
  return PC_OK;
}

int vtu_write(const char* path, int n, const pc_vtu_patch* pt) {
  if (!path) return fail(PC_E_INVALID, "vtu: invalid argument");
  CellGrid g;
  CHK(collect_cells(n, pt, 0.0, g, "vtu"));
  const std::vector<double>& pts = g.pts;
  const std::vector<int32_t>& conn = g.conn;
  const int kArrays = kVtuArrays;
  const char* const* names = kVtuNames;
  const std::vector<double>* cell = g.cell;
  const size_t nv = pts.size() / 3, nc = conn.size() / 4;
  std::vector<int32_t> offs(nc);
  for (size_t c = 0; c < nc; ++c) offs[c] = (int32_t)(4 * (c + 1));
  std::vector<uint8_t> types(nc, 9);  // VTK_QUAD
  // appended blocks: UInt32 byte count, then the bytes
  struct Block {
    const void* data;
    size_t bytes;
  };
  std::vector<Block> blocks;
  blocks.push_back({pts.data(), pts.size() * sizeof(double)});
  blocks.push_back({conn.data(), conn.size() * sizeof(int32_t)});
  blocks.push_back({offs.data(), offs.size() * sizeof(int32_t)});
  blocks.push_back({types.data(), types.size()});
  for (int a = 0; a < kArrays; ++a) blocks.push_back({cell[a].data(), cell[a].size() * sizeof(double)});
  for (const Block& b : blocks)
    if (b.bytes > 0xffffffffu) return fail(PC_E_INVALID, "vtu: array larger than 4 GiB");
  std::vector<size_t> off(blocks.size());
  size_t o = 0;
  for (size_t k = 0; k < blocks.size(); ++k) {
    off[k] = o;
    o += 4 + blocks[k].bytes;
  }
  std::string x;
  x += "<?xml version=\"1.0\"?>\n";
  x += "<VTKFile type=\"UnstructuredGrid\" version=\"0.1\" byte_order=\"LittleEndian\" "
       "header_type=\"UInt32\">\n";
  x += "<UnstructuredGrid>\n";
  x += "<Piece NumberOfPoints=\"" + std::to_string(nv) + "\" NumberOfCells=\"" +
       std::to_string(nc) + "\">\n";
  auto arr = [&](const char* type, const char* name, int comps, size_t k) {
    x += std::string("<DataArray type=\"") + type + "\" Name=\"" + name +
         "\" NumberOfComponents=\"" + std::to_string(comps) +
         "\" format=\"appended\" offset=\"" + std::to_string(off[k]) + "\"/>\n";
  };
  x += "<Points>\n";
  arr("Float64", "points", 3, 0);
  x += "</Points>\n<Cells>\n";
  arr("Int32", "connectivity", 1, 1);
  arr("Int32", "offsets", 1, 2);
  arr("UInt8", "types", 1, 3);
  x += "</Cells>\n<CellData>\n";
  for (int a = 0; a < kArrays; ++a) arr("Float64", names[a], 1, 4 + a);
  x += "</CellData>\n</Piece>\n</UnstructuredGrid>\n<AppendedData encoding=\"raw\">\n_";
  FILE* f = std::fopen(path, "wb");
  if (!f) return fail(PC_E_INVALID, std::string("vtu: cannot open ") + path);
  bool ok = std::fwrite(x.data(), 1, x.size(), f) == x.size();
  for (const Block& b : blocks) {
    const uint32_t len = (uint32_t)b.bytes;
    ok = ok && std::fwrite(&len, 4, 1, f) == 1;
    if (b.bytes) ok = ok && std::fwrite(b.data, 1, b.bytes, f) == b.bytes;
  }
  const char* tail = "\n</AppendedData>\n</VTKFile>\n";
  ok = ok && std::fwrite(tail, 1, std::strlen(tail), f) == std::strlen(tail);
  ok = (std::fclose(f) == 0) && ok;
  if (!ok) return fail(PC_E_INVALID, std::string("vtu: write failed for ") + path);
  return PC_OK;
}

// Legacy VTK (ASCII) of the same cells, the format of the reference's
// VTKSubgridPlotter (grid/plotter/VTKSubgridPlotter.cpp:9-241, through
// Peano's VTK text writer): POINTS, CELLS (quads), CELL_TYPES 9, and one
// SCALARS block per cell array; values printed with 17 significant digits
// (exact round trip).
int vtk_write(const char* path, int n, const pc_vtu_patch* pt) {
  if (!path) return fail(PC_E_INVALID, "vtk: invalid argument");
  CellGrid g;
  CHK(collect_cells(n, pt, 0.015, g, "vtk"));
  const size_t nv = g.pts.size() / 3, nc = g.conn.size() / 4;
  FILE* f = std::fopen(path, "w");
  if (!f) return fail(PC_E_INVALID, std::string("vtk: cannot open ") + path);
  bool ok = std::fprintf(f, "# vtk DataFile Version 2.0\nPeanoClaw subgrids\nASCII\n"
                            "DATASET UNSTRUCTURED_GRID\nPOINTS %zu double\n", nv) > 0;
  for (size_t v = 0; v < nv && ok; ++v)
    ok = std::fprintf(f, "%.17g %.17g %.17g\n", g.pts[3 * v], g.pts[3 * v + 1],
                      g.pts[3 * v + 2]) > 0;
  ok = ok && std::fprintf(f, "CELLS %zu %zu\n", nc, 5 * nc) > 0;
  for (size_t c = 0; c < nc && ok; ++c)
    ok = std::fprintf(f, "4 %d %d %d %d\n", g.conn[4 * c], g.conn[4 * c + 1], g.conn[4 * c + 2],
                      g.conn[4 * c + 3]) > 0;
  ok = ok && std::fprintf(f, "CELL_TYPES %zu\n", nc) > 0;
  for (size_t c = 0; c < nc && ok; ++c) ok = std::fputs("9\n", f) >= 0;
  ok = ok && std::fprintf(f, "CELL_DATA %zu\n", nc) > 0;
  for (int a = 0; a < kVtuArrays && ok; ++a) {
    ok = std::fprintf(f, "SCALARS %s double 1\nLOOKUP_TABLE default\n", kVtuNames[a]) > 0;
    for (size_t c = 0; c < nc && ok; ++c) ok = std::fprintf(f, "%.17g\n", g.cell[a][c]) > 0;
  }
  ok = (std::fclose(f) == 0) && ok;
  if (!ok) return fail(PC_E_INVALID, std::string("vtk: write failed for ") + path);
  return PC_OK;
}

// The listed slots of a pool as plotter patches (uNew of the current state,
// b and the descriptor; age 0); `data` keeps the downloaded arrays.
int pool_patches(pc_pool* p, const int32_t* slots, int nslots,
                 std::vector<std::vector<double>>& data, std::vector<pc_vtu_patch>& pt) {
  if (!p || nslots < 0 || (nslots > 0 && !slots)) return fail(PC_E_INVALID, "invalid argument");
  data.assign((size_t)nslots, {});
  pt.assign((size_t)nslots, pc_vtu_patch{});
  for (int k = 0; k < nslots; ++k) {
    const int s = slots[k];
    if (s < 0 || s >= (int)p->slots.size() || !p->slots[s].live)
      return fail(PC_E_INVALID, "plotter: slot is not live");
    pc_cell_description d;
    CHK(pc_patch_get_description(p, s, &d));
    const int nx = d.subdivision_factor[0], ny = d.subdivision_factor[1];
    const size_t a = (size_t)nx * ny;
    data[k].assign(4 * a, 0.0);
    double* unew = data[k].data();
    double* b = unew + 3 * a;
    std::vector<double> bw((size_t)(nx + 2) * (ny + 2));
    CHK(pc_patch_download(p, s, unew, nullptr, bw.data()));
    for (int i = 0; i < nx; ++i)
      for (int j = 0; j < ny; ++j) b[(size_t)i * ny + j] = bw[(size_t)(i + 1) * (ny + 2) + j + 1];
    pc_vtu_patch& P = pt[k];
    P.subdivision_factor[0] = nx;
    P.subdivision_factor[1] = ny;
    P.ghostlayer_width = d.ghostlayer_width;
    P.level = d.level;
    P.position[0] = d.position[0];
    P.position[1] = d.position[1];
    P.size[0] = d.size[0];
    P.size[1] = d.size[1];
    P.time = d.time;
    P.timestep_size = d.timestep_size;
    P.demanded_mesh_width = (std::fabs(d.demanded_mesh_width[0]) +
                             std::fabs(d.demanded_mesh_width[1])) / 2.0;
    P.age = 0.0;
    for (int q = 0; q < 3; ++q) P.q[q] = unew + (size_t)q * a;
    P.b = b;
  }
  return PC_OK;
}

}  // namespace

extern "C" {

int pc_vtu_write(const char* path, int npatches, const pc_vtu_patch* patches) {
  return vtu_write(path, npatches, patches);
}

int pc_pool_write_vtu(pc_pool* p, const char* path, const int32_t* slots, int nslots) {
  if (!path) return fail(PC_E_INVALID, "invalid argument");
  std::vector<std::vector<double>> data;
  std::vector<pc_vtu_patch> pt;
  CHK(pool_patches(p, slots, nslots, data, pt));
  return vtu_write(path, nslots, pt.data());
}

int pc_vtk_write(const char* path, int npatches, const pc_vtu_patch* patches) {
  return vtk_write(path, npatches, patches);
}

int pc_pool_write_vtk(pc_pool* p, const char* path, const int32_t* slots, int nslots) {
  if (!path) return fail(PC_E_INVALID, "invalid argument");
  std::vector<std::vector<double>> data;
  std::vector<pc_vtu_patch> pt;
  CHK(pool_patches(p, slots, nslots, data, pt));
  return vtk_write(path, nslots, pt.data());
}

}  // extern "C"
