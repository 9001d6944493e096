// Sphere plan of one structure: the per-site point lists and projector
// tables (pp_sphere.hip) and the line-block work lists of the fused x pass
// with sphere sums (pp_fused.hip).
#pragma once
#include <vector>

#include "pp_internal.h"

namespace pp {

// per-site geometry handed to the table kernels
struct SiteGeo {
  double coord[3];
  double rmax;
  double R0;
  int g[3];
  int c[3];
  int elem;
  int nproj;
  int nchan;
  int ngrid;
  long spl_off;  // offset (doubles) of the species spline block
  int chan_off;  // offset into the channel-l array
};

// R0, rmax, ngrid and the candidate box (g, c) of a site whose coord is
// set (pp_sphere.hip).
void site_box(const double *lattice, const int *fftg, double rmax, int ngrid, SiteGeo &g);

// Fused x pass: a work item is a block of FX_BY x FX_BZ x-lines (y, z) and
// up to FX_G sites whose spheres reach them; each of the four wavefronts
// of a work-group accumulates FX_S of those sites.
constexpr int FX_BY = 4, FX_BZ = 4, FX_BL = FX_BY * FX_BZ;
constexpr int FX_S = 3, FX_G = 4 * FX_S;
constexpr int FX_PMAX = 18;  // projectors per site the fused kernel holds

// Work lists of the fused x pass (host).  items: (y0, z0, sites, 0) per
// item; isites: FX_G slots per item (site or -1); bitem: items of block b
// are bitem[b] .. bitem[b+1]-1 (block b = (y / FX_BY) * nbz + z / FX_BZ);
// rows_off / rows: the partial rows item * FX_G + slot of every site.
struct FusedLists {
  std::vector<int> items, isites, bitem, rows_off, rows;
  int nitems = 0, nblk = 0, nbz = 0;
};
// false when a site's box reaches a (y, z) line twice (2 g + 1 > N in y or z)
bool fused_lists(const double *lattice, const int *fftg, const std::vector<SiteGeo> &geo, FusedLists &fl);

}  // namespace pp

struct pp_sphere {
  pp_ctx *ctx = nullptr;
  int precision = PP_FP64;
  double lattice[9];
  int fftg[3];
  int nsites = 0, nelems = 0;
  std::vector<int> labels;
  std::vector<double> frac;
  std::vector<int> nchan, ls_off, ngrid, grid_off;
  std::vector<int> ls;
  std::vector<double> rmax;
  std::vector<int> nproj_site, proj_off;  // host
  std::vector<int> site_pt_h;
  int ptot = 0, pmax = 0, pmin = 0;
  long npts = 0;
  long gen = 0;  // build count: graphs captured against older tables refuse to replay
  // device
  pp::DevBuf d_spline;  // per species: x, f, b, c, d [nchan][ngrid]
  pp::DevBuf d_sitemeta;
  pp::DevBuf d_site_pt, d_site_np, d_tab_off, d_proj_off, d_counts;
  pp::DevBuf d_proj_m;  // m of every projector, [ptot]
  pp::DevBuf d_pidx, d_path, d_table;
  pp::DevBuf d_chan;  // per species channel l list
  // Segments: maximal runs of consecutive x of one site's points on one
  // (y, z) line, as int4 (first point, length, wrapped x of the first
  // point, site) and the wrapped line y * N2 + z.
  long nseg = 0;
  std::vector<int> seg_off_h;
  pp::DevBuf d_seg, d_segline, d_seg_off;
  // per x-line (y * N2 + z) bit mask of the points some sphere holds,
  // xmask_words 32-bit words per line
  pp::DevBuf d_xmask;
  int xmask_words = 0;
  // fused x pass work lists (pp_fused.hip); fused_ok is false when the plan
  // does not fit them (FP64 tables, more than FX_PMAX projectors on a site,
  // or a site box that wraps onto itself in y or z)
  bool fused_ok = false;
  bool fx_built = false;          // work lists built for the current tables
  std::vector<pp::SiteGeo> geo_h;  // site geometry of the current tables
  int fx_nitems = 0, fx_nrows = 0, fx_nrows_used = 0, fx_nblocks = 0, fx_nbz = 0;
  int fx_maxlen = 1;  // longest possible segment (widest site box in x)
  std::vector<int> fx_host;  // packed host copy of the work lists
  pp::DevBuf d_fx;           // items (int4), item sites, block -> items, site -> rows
  long fx_items_off = 0, fx_sites_off = 0, fx_bitem_off = 0, fx_rows_off = 0, fx_rows = 0;
  pp::DevBuf d_segtab;  // [item][line][slot] segment id or -1
  pp::DevBuf d_fxflag;  // set when a segment finds no slot (never, by construction)
};
