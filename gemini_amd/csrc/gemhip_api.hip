// gemhip_api.hip -- C ABI of the transport library (declared in include/gemini3d_hip.h).
//
// Holds the single library context (the reference keeps sizes and the mesh in module
// globals and Fortran objects; here they are registered once), moves state between the
// caller's host blocks and HBM according to the residency mode, and implements the
// reference-named entry points on top of the kernel launchers.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <string>
#include <vector>

#include "../../include/gemini3d_hip.h"
#include "gemhip_internal.h"
#include "gemhip_io.h"

// Work space of gemhip_energy_diffusion_noalloc (TRBDF2 pivot rows, right-hand sides,
// solutions and row interchanges): a static device array of NOALLOC_SCRATCH_BYTES, one copy
// on every GPU the library runs on, held until the process ends.  Host code never uses
// this name as an address: gemhip_init looks the array's device address up once per
// device (noalloc_scratch below) and the kernels receive that address.
#define NOALLOC_SCRATCH_BYTES (256ul << 20)
static __device__ __align__(256) unsigned char scratch[NOALLOC_SCRATCH_BYTES];

namespace gemhip {
namespace {

// device address of `scratch` on each device (index: device ordinal), set by gemhip_init
constexpr int MAX_DEVICES = 64;
void* noalloc_scratch[MAX_DEVICES] = {};
hipError_t noalloc_lookup[MAX_DEVICES] = {};   // result of the address lookup, per device
bool noalloc_looked_up[MAX_DEVICES] = {};

struct TimedLaunch {
  std::string name;
  hipEvent_t a, b;
};

struct Ctx {
  bool inited = false;
  int device = 0;
  hipStream_t stream = nullptr;
  int lx1 = 0, lx2 = 0, lx3 = 0;
  bool sized = false, grid_ok = false, neutrals_ok = false;
  int xtype = 0;
  int diffsolvetype = 2;
  double Teinf = 1500.0;
  DevGrid g{};
  DevState s{};
  double* d_geom = nullptr;
  double* d_fluid = nullptr;
  double* d_aux = nullptr;
  double* d_J1 = nullptr;        // J1 (electrovars field 4), then E1 (field 1): 2 fields
  double* d_vsi = nullptr;
  double* d_v1b = nullptr;
  double* d_neut = nullptr;
  double* d_scratch = nullptr;
  size_t scratch_bytes = 0;
  double* d_Q = nullptr;        // artificial viscosity (intvars%Q), lx1*lx2*lx3*7
  double* d_src = nullptr;      // source/loss inputs (src_layout), allocated on first use
  bool atmos_ok = false;        // gemhip_register_atmosphere called
  bool gravity_ok = false;      // gemhip_register_gravity called
  double* d_rates = nullptr;    // dt_calc partial maxima and results
  double tcfl = 0.9;            // cfg%tcfl
  int flagglow = 0;             // cfg%flagglow
  int dt_calls = 0;             // the reference's iteration count as dt_select sees it
  int* d_null = nullptr;        // null-cell mask, lx1*lx2*lx3 (null when none registered)
  std::string outdir;           // cfg%outdir, cfg%flagoutput and the file compression level
  int flagoutput = 1;
  int comp_lvl = 3;
  bool output_ok = false;       // gemhip_register_output called
  double* d_out = nullptr;      // output staging: J1 J2 J3 Phi, then the packed real32 datasets
  size_t out_bytes = 0;
  float* h_out = nullptr;       // pinned host copy of the packed datasets
  size_t h_out_bytes = 0;
  double* d_kappa = nullptr;    // fitting extension (lx2*lx3 each)
  double* d_qtop = nullptr;
  double* d_adj = nullptr;      // adjoint work: a second field and the column scratch
  size_t adj_bytes = 0;
  bool vsi_current = true;      // full intvars%vs*i materialised
  int residency = 0;
  const double* pushed_fluid = nullptr;
  int error_mode = 0;
  std::string last_error;
  Tuning tune;
  bool timing = false;
  std::vector<TimedLaunch> pending;
  std::vector<hipEvent_t> pool;
  std::map<std::string, std::pair<double, long>> acc;
  std::map<std::string, hipEvent_t> open;
};

// The reference API acts on one implicit context (handle 0).  Ensembles create more
// (gemhip_context_create) and select one at a time; every call acts on the selected one.
Ctx& default_ctx() {
  static Ctx c;
  return c;
}
std::vector<Ctx*>& contexts() {          // handle h >= 1 is contexts()[h - 1] (null once freed)
  static std::vector<Ctx*> v;
  return v;
}
Ctx* g_current = nullptr;
int g_current_handle = 0;

Ctx& ctx() { return g_current ? *g_current : default_ctx(); }

void fail(const char* fmt, ...) {
  char buf[1024];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  Ctx& c = ctx();
  if (c.error_mode == 0) {
    fprintf(stderr, "gemini3d_hip: %s\n", buf);
    fflush(stderr);
    abort();
  }
  if (c.last_error.empty()) c.last_error = buf;
}

bool hipok(hipError_t e, const char* what) {
  if (e == hipSuccess) return true;
  fail("%s: %s", what, hipGetErrorString(e));
  return false;
}

hipEvent_t take_event() {
  Ctx& c = ctx();
  if (!c.pool.empty()) {
    hipEvent_t e = c.pool.back();
    c.pool.pop_back();
    return e;
  }
  hipEvent_t e;
  hipEventCreate(&e);
  return e;
}

void collect_timings() {
  Ctx& c = ctx();
  for (auto& t : c.pending) {
    hipEventSynchronize(t.b);
    float ms = 0.f;
    hipEventElapsedTime(&ms, t.a, t.b);
    auto& a = c.acc[t.name];
    a.first += ms;
    a.second += 1;
    c.pool.push_back(t.a);
    c.pool.push_back(t.b);
  }
  c.pending.clear();
}

void free_all() {
  Ctx& c = ctx();
  for (double** p : {&c.d_geom, &c.d_fluid, &c.d_aux, &c.d_J1, &c.d_vsi, &c.d_v1b, &c.d_neut,
                     &c.d_scratch, &c.d_kappa, &c.d_qtop, &c.d_adj, &c.d_Q, &c.d_rates, &c.d_src}) {
    if (*p) hipFree(*p);
    *p = nullptr;
  }
  if (c.d_null) hipFree(c.d_null);
  c.d_null = nullptr;
  if (c.d_out) hipFree(c.d_out);
  if (c.h_out) hipHostFree(c.h_out);
  c.d_out = nullptr;
  c.h_out = nullptr;
  c.out_bytes = c.h_out_bytes = 0;
  c.scratch_bytes = 0;
  c.adj_bytes = 0;
  c.s.kappa = c.s.qtop = nullptr;
  c.sized = c.grid_ok = c.neutrals_ok = false;
  c.atmos_ok = c.gravity_ok = false;
}

bool dev_alloc(double** p, size_t n, const char* what) {
  if (!hipok(hipMalloc((void**)p, n * sizeof(double)), what)) return false;
  return hipok(hipMemset(*p, 0, n * sizeof(double)), what);
}

long nv1i() { Ctx& c = ctx(); return (long)(c.lx1 + 1) * c.lx2 * c.lx3; }
long nv2i() { Ctx& c = ctx(); return (long)c.lx1 * (c.lx2 + 1) * c.lx3; }
long nv3i() { Ctx& c = ctx(); return (long)c.lx1 * c.lx2 * (c.lx3 + 1); }

bool ready(bool need_neutrals = false) {
  Ctx& c = ctx();
  if (!c.inited) { fail("library not initialised (gemhip_init)"); return false; }
  if (!c.sized) { fail("grid size not set (gemhip_set_grid_size)"); return false; }
  if (!c.grid_ok) { fail("grid not registered (gemhip_register_grid)"); return false; }
  if (need_neutrals && !c.neutrals_ok) {
    fail("neutral atmosphere not registered (gemhip_register_neutrals)");
    return false;
  }
  return true;
}

// libgemini_c.f90:63-82: only Cartesian (1) and dipole (2) meshes exist
bool check_xtype(const int* xtype) {
  Ctx& c = ctx();
  if (!xtype || (*xtype != 1 && *xtype != 2)) {
    fail("Unable to identify object type during conversion from C to Fortran class pointer:  %d",
         xtype ? *xtype : -1);
    return false;
  }
  if (*xtype != c.xtype) {
    fail("mesh type %d differs from the registered mesh type %d", *xtype, c.xtype);
    return false;
  }
  return true;
}

// ---- host <-> HBM movement of whole variables (7 species fields each) ----
enum Var { NS = 0, VS1 = 1, VS2 = 2, VS3 = 3, TS = 4, RHOVS1 = 5, RHOES = 6, J1 = 7, E1 = 8 };

double* dev_var(Var v) {
  Ctx& c = ctx();
  const long n7 = 7L * c.g.N;
  switch (v) {
    case NS: return c.d_fluid;
    case VS1: return c.d_fluid + n7;
    case VS2: return c.d_fluid + 2 * n7;
    case VS3: return c.d_fluid + 3 * n7;
    case TS: return c.d_fluid + 4 * n7;
    case RHOVS1: return c.d_aux;
    case RHOES: return c.d_aux + n7;
    case E1: return c.d_J1 + c.g.N;
    default: return c.d_J1;
  }
}

const double* host_var(Var v, double* const* fluidvars, double* const* aux, double* const* ev) {
  Ctx& c = ctx();
  const long n7 = 7L * c.g.N;
  switch (v) {
    case NS: case VS1: case VS2: case VS3: case TS: return *fluidvars + (long)v * n7;
    case RHOVS1: return *aux;
    case RHOES: return *aux + n7;
    case E1: return *ev;
    default: return *ev + 3L * c.g.N;
  }
}

bool check_host(double* const* p, const char* name) {
  if (!p || !*p) { fail("%s is a null pointer", name); return false; }
  return true;
}

struct Io {
  double* const* fv;
  double* const* av;
  double* const* ev;
};

bool upload(const Io& io, std::initializer_list<Var> vars) {
  Ctx& c = ctx();
  if (c.residency == 1) {
    if (io.fv && *io.fv && c.pushed_fluid && *io.fv != c.pushed_fluid) {
      fail("fluidvars differs from the block pushed with gemhip_state_push");
      return false;
    }
    return true;
  }
  for (Var v : vars) {
    const size_t n = (v == J1 || v == E1 ? 1L : 7L) * c.g.N;
    if (!hipok(hipMemcpyAsync(dev_var(v), host_var(v, io.fv, io.av, io.ev), n * sizeof(double),
                              hipMemcpyDefault, c.stream), "upload"))
      return false;
  }
  return true;
}

bool download(const Io& io, std::initializer_list<Var> vars) {
  Ctx& c = ctx();
  if (c.residency == 1) return true;
  for (Var v : vars) {
    const size_t n = 7L * c.g.N;
    double* h = const_cast<double*>(host_var(v, io.fv, io.av, io.ev));
    if (!hipok(hipMemcpyAsync(h, dev_var(v), n * sizeof(double), hipMemcpyDefault, c.stream),
               "download"))
      return false;
  }
  return hipok(hipStreamSynchronize(c.stream), "synchronise");
}

void sweep_entry(double** fv, double** av, int* xtype, double* dt, int dir, int fields) {
  if (!ready() || !check_xtype(xtype)) return;
  if (!dt) { fail("dt is a null pointer"); return; }
  Ctx& c = ctx();
  if (!c.vsi_current) {
    fail("interface velocities are not materialised; call interface_vels_allspec_C first");
    return;
  }
  Io io{fv, av, nullptr};
  std::vector<Var> vars;
  if (fields & F_MASS) vars.push_back(NS);
  if (fields & F_MOMENTUM) vars.push_back(RHOVS1);
  if (fields & F_ENERGY) vars.push_back(RHOES);
  if ((fields & F_MASS) && !check_host(fv, "fluidvars")) return;
  if ((fields & (F_MOMENTUM | F_ENERGY)) && !check_host(av, "fluidauxvars")) return;
  for (Var v : vars) if (!upload(io, {v})) return;
  launch_sweep(c.g, c.s, dir, fields, *dt, false, c.stream, c.tune);
  if (!hipok(hipGetLastError(), "sweep launch")) return;
  for (Var v : vars) if (!download(io, {v})) return;
}

bool ensure_scratch() {
  Ctx& c = ctx();
  const size_t need = diffusion_scratch_bytes(c.g);
  if (c.d_scratch && c.scratch_bytes >= need) return true;
  if (c.d_scratch) hipFree(c.d_scratch);
  c.d_scratch = nullptr;
  if (!hipok(hipMalloc((void**)&c.d_scratch, need), "scratch allocation")) return false;
  c.scratch_bytes = need;
  return true;
}

bool run_energy_diffusion(double dt, bool floor_ghosts = true) {
  Ctx& c = ctx();
  if (c.diffsolvetype != 1 && c.diffsolvetype != 2) {
    fail("Unsupported diffusion solver type/mode:  %d.  Should be either 1 or 2.",
         c.diffsolvetype);
    return false;
  }
  if (c.lx1 < 4) { fail("energy diffusion needs lx1 >= 4 (got %d)", c.lx1); return false; }
  if ((c.s.kappa || c.s.qtop) && c.diffsolvetype != 2) {
    fail("the conductivity scale and top heat flux need the TRBDF2 solver (diffsolvetype 2)");
    return false;
  }
  if (c.s.qtop && c.g.gridflag != 2) {
    fail("a top heat flux needs gridflag 2 (top at x1 = lx1), got %d", c.g.gridflag);
    return false;
  }
  if (!ensure_scratch()) return false;
  launch_energy_diffusion(c.g, c.s, dt, c.diffsolvetype, c.Teinf, c.d_scratch, c.stream, c.tune,
                          floor_ghosts);
  return hipok(hipGetLastError(), "energy diffusion launch");
}

void halo_fields(std::initializer_list<Var> vars) {
  Ctx& c = ctx();
  double* f[8];
  int n = 0;
  for (Var v : vars) f[n++] = dev_var(v);
  launch_halo_periodic(c.g, f, n, 2, LSP, c.stream);
}

// device evaluation of the scalar routines the kernels use, for the math checks
__global__ void k_eval_math(int op, const double* x, double* y, long n) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  if (op == 4) {                                    // pairs (numerator, divisor)
    const double a = x[2 * i], b = x[2 * i + 1];
    y[i] = div_rcp(a, b, 1.0 / b);
    return;
  }
  const double v = x[i];
  y[i] = op == 0 ? pow25(v) : op == 1 ? sqrt(v) : op == 2 ? 1.0 / v : pow(v, 2.5);
}

}  // namespace

void timing_begin(const char* name, hipStream_t st) {
  Ctx& c = ctx();
  if (!c.timing) return;
  hipEvent_t e = take_event();
  hipEventRecord(e, st);
  c.open[name] = e;
}

void timing_end(const char* name, hipStream_t st) {
  Ctx& c = ctx();
  if (!c.timing) return;
  auto it = c.open.find(name);
  if (it == c.open.end()) return;
  hipEvent_t e = take_event();
  hipEventRecord(e, st);
  c.pending.push_back({name, it->second, e});
  c.open.erase(it);
  if (c.pending.size() > 4096) collect_timings();
}

}  // namespace gemhip

using namespace gemhip;

extern "C" {

int gemhip_init(int device) {
  Ctx& c = ctx();
  if (c.inited) {
    if (device == c.device) return 0;
    fail("already initialised on device %d", c.device);
    return 1;
  }
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess || n <= 0) { fail("no HIP device available"); return 2; }
  if (device < 0 || device >= n) { fail("device %d out of range (%d devices)", device, n); return 3; }
  if (!hipok(hipSetDevice(device), "hipSetDevice")) return 4;
  if (!hipok(hipStreamCreateWithFlags(&c.stream, hipStreamNonBlocking), "stream")) return 5;
  // the allocation-free entry point's work space on this device: its address, looked up
  // once per device with the device current (the result is checked where it is used)
  if (device < MAX_DEVICES && !noalloc_looked_up[device]) {
    void* p = nullptr;
    noalloc_lookup[device] = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    noalloc_scratch[device] = noalloc_lookup[device] == hipSuccess ? p : nullptr;
    noalloc_looked_up[device] = true;
  }
  c.device = device;
  c.inited = true;
  return 0;
}

void gemhip_finalize(void) {
  Ctx& c = ctx();
  if (!c.inited) return;
  hipStreamSynchronize(c.stream);
  collect_timings();
  for (auto e : c.pool) hipEventDestroy(e);
  c.pool.clear();
  free_all();
  hipStreamDestroy(c.stream);
  c.stream = nullptr;
  c.inited = false;
}

void gemhip_set_error_mode(int mode) { ctx().error_mode = mode ? 1 : 0; }
const char* gemhip_last_error(void) { return ctx().last_error.c_str(); }
void gemhip_clear_error(void) { ctx().last_error.clear(); }

int gemhip_set_grid_size(int lx1, int lx2, int lx3) {
  Ctx& c = ctx();
  if (!c.inited) { fail("library not initialised (gemhip_init)"); return 1; }
  // a field's cell index fits in 32 bits (2^31 ghosted cells are 16 GiB per field, far
  // more than the 35 fields of a state can hold in one GPU's memory)
  if (lx1 < 4 || lx2 < 1 || lx3 < 1 ||
      (long)(lx1 + 4) * (lx2 + 4) * (lx3 + 4) >= (1L << 31)) {
    fail("grid size %d x %d x %d not supported (lx1 >= 4, lx2, lx3 >= 1, fewer than 2^31 "
         "ghosted cells)", lx1, lx2, lx3);
    return 2;
  }
  hipStreamSynchronize(c.stream);
  free_all();
  c.lx1 = lx1; c.lx2 = lx2; c.lx3 = lx3;
  DevGrid& g = c.g;
  g = DevGrid{};
  g.lx1 = lx1; g.lx2 = lx2; g.lx3 = lx3;
  g.L1 = lx1 + 4; g.L2 = lx2 + 4; g.L3 = lx3 + 4;
  g.N = (long)g.L1 * g.L2 * g.L3;
  const long ni = (long)lx1 * lx2 * lx3;
  bool ok = dev_alloc(&c.d_fluid, 35L * g.N, "fluidvars") &&
            dev_alloc(&c.d_aux, 14L * g.N, "fluidauxvars") &&
            dev_alloc(&c.d_J1, 2 * (size_t)g.N, "J1 and E1") &&
            dev_alloc(&c.d_vsi, 7L * (nv1i() + nv2i() + nv3i()), "interface velocities") &&
            dev_alloc(&c.d_v1b, 14L * lx2 * lx3, "x1 boundary velocities") &&
            dev_alloc(&c.d_neut, 4L * ni, "neutrals") &&
            dev_alloc(&c.d_Q, 7L * ni, "artificial viscosity");
  if (!ok) { free_all(); return 3; }
  DevState& s = c.s;
  s = DevState{};
  s.ns = dev_var(NS); s.vs1 = dev_var(VS1); s.vs2 = dev_var(VS2); s.vs3 = dev_var(VS3);
  s.Ts = dev_var(TS); s.rhovs1 = dev_var(RHOVS1); s.rhoes = dev_var(RHOES); s.J1 = c.d_J1;
  s.vs1i = c.d_vsi; s.vs2i = s.vs1i + 7L * nv1i(); s.vs3i = s.vs2i + 7L * nv2i();
  s.v1b = c.d_v1b;
  s.nn = c.d_neut; s.Tn = c.d_neut + 3 * ni;
  c.sized = true;
  c.vsi_current = true;
  c.pushed_fluid = nullptr;
  return 0;
}

int gemhip_register_grid(int xtype, int gridflag, int flagper, const int* isglobal6,
                         const double* h1, const double* h2, const double* h3,
                         const double* h1x1i, const double* h2x1i, const double* h3x1i,
                         const double* h1x2i, const double* h2x2i, const double* h3x2i,
                         const double* h1x3i, const double* h2x3i, const double* h3x3i,
                         const double* dx1, const double* dx1i, const double* dx2,
                         const double* dx2i, const double* dx3, const double* dx3i) {
  Ctx& c = ctx();
  if (!c.sized) { fail("grid size not set (gemhip_set_grid_size)"); return 1; }
  if (xtype != 1 && xtype != 2) {
    fail("Unable to identify object type during conversion from C to Fortran class pointer:  %d",
         xtype);
    return 2;
  }
  if (gridflag < 0 || gridflag > 2) { fail("gridflag %d not in 0..2", gridflag); return 3; }
  const double* src[] = {h1, h2, h3, h1x1i, h2x1i, h3x1i, h1x2i, h2x2i, h3x2i,
                         h1x3i, h2x3i, h3x3i, dx1, dx1i, dx2, dx2i, dx3, dx3i};
  for (const double* p : src)
    if (!p) { fail("null geometry pointer"); return 4; }
  if (!isglobal6) { fail("null isglobal pointer"); return 4; }
  DevGrid& g = c.g;
  const long n[] = {g.N, g.N, g.N, nv1i(), nv1i(), nv1i(), nv2i(), nv2i(), nv2i(),
                    nv3i(), nv3i(), nv3i(), c.lx1 + 3, c.lx1, c.lx2 + 3, c.lx2, c.lx3 + 3, c.lx3};
  long total = 0;
  for (long k : n) total += k;
  g.hstride = diffusion_hcol_doubles(g) / 3;
  if (c.d_geom) hipFree(c.d_geom);
  c.d_geom = nullptr;
  // every array starts 256-byte aligned (the sweeps' DMA loads read 16-byte pairs)
  if (!hipok(hipMalloc((void**)&c.d_geom, (total + 19 * 32 + 3 * g.hstride) * sizeof(double)),
             "geometry"))
    return 5;
  const double** dst[] = {&g.h1, &g.h2, &g.h3, &g.h1x1i, &g.h2x1i, &g.h3x1i, &g.h1x2i, &g.h2x2i,
                          &g.h3x2i, &g.h1x3i, &g.h2x3i, &g.h3x3i, &g.dx1, &g.dx1i, &g.dx2, &g.dx2i,
                          &g.dx3, &g.dx3i};
  long off = 0;
  for (int k = 0; k < 18; ++k) {
    *dst[k] = c.d_geom + off;
    if (!hipok(hipMemcpy(c.d_geom + off, src[k], n[k] * sizeof(double), hipMemcpyDefault),
               "geometry upload"))
      return 6;
    off = (off + n[k] + 31) & ~31L;
  }
  off = (off + 31) & ~31L;                     // 256-byte aligned rows of 64 columns
  g.hcol = c.d_geom + off;                     // the diffusion's copy of h1, h2, h3
  launch_fill_hcol(g, c.d_geom + off, c.stream);
  if (!hipok(hipStreamSynchronize(c.stream), "geometry reorder")) return 6;
  g.gridflag = gridflag;
  g.flagper = flagper ? 1 : 0;
  for (int k = 0; k < 6; ++k) g.isglobal[k] = isglobal6[k] ? 1 : 0;
  c.xtype = xtype;
  c.grid_ok = true;
  return 0;
}

int gemhip_register_neutrals(const double* nn, const double* Tn) {
  Ctx& c = ctx();
  if (!c.sized) { fail("grid size not set (gemhip_set_grid_size)"); return 1; }
  if (!nn || !Tn) { fail("null neutral pointer"); return 2; }
  const long ni = (long)c.lx1 * c.lx2 * c.lx3;
  if (!hipok(hipMemcpy(c.d_neut, nn, 3 * ni * sizeof(double), hipMemcpyDefault), "nn upload") ||
      !hipok(hipMemcpy(c.d_neut + 3 * ni, Tn, ni * sizeof(double), hipMemcpyDefault),
             "Tn upload"))
    return 3;
  c.neutrals_ok = true;
  return 0;
}

int gemhip_register_config(int diffsolvetype, double Teinf) {
  Ctx& c = ctx();
  c.diffsolvetype = diffsolvetype;
  c.Teinf = Teinf;
  return 0;
}

int gemhip_set_residency(int mode) {
  Ctx& c = ctx();
  if (mode != 0 && mode != 1) { fail("residency mode %d not 0 or 1", mode); return 1; }
  if (c.inited) hipStreamSynchronize(c.stream);
  c.residency = mode;
  return 0;
}

int gemhip_state_push(const double* fluidvars, const double* fluidauxvars,
                      const double* electrovars) {
  if (!ready()) return 1;
  Ctx& c = ctx();
  if (!fluidvars) { fail("fluidvars is a null pointer"); return 2; }
  const size_t N = (size_t)c.g.N;
  bool ok = hipok(hipMemcpyAsync(c.d_fluid, fluidvars, 35 * N * sizeof(double),
                                 hipMemcpyDefault, c.stream), "push fluidvars");
  if (ok && fluidauxvars)
    ok = hipok(hipMemcpyAsync(c.d_aux, fluidauxvars, 14 * N * sizeof(double),
                              hipMemcpyDefault, c.stream), "push fluidauxvars");
  if (ok && electrovars)
    ok = hipok(hipMemcpyAsync(c.d_J1, electrovars + 3 * N, N * sizeof(double),
                              hipMemcpyDefault, c.stream), "push J1") &&
         hipok(hipMemcpyAsync(c.d_J1 + N, electrovars, N * sizeof(double), hipMemcpyDefault,
                              c.stream), "push E1");
  if (ok) ok = hipok(hipStreamSynchronize(c.stream), "push");
  c.pushed_fluid = fluidvars;
  return ok ? 0 : 3;
}

int gemhip_state_pull(double* fluidvars, double* fluidauxvars) {
  if (!ready()) return 1;
  Ctx& c = ctx();
  const size_t N = (size_t)c.g.N;
  bool ok = true;
  if (fluidvars)
    ok = hipok(hipMemcpyAsync(fluidvars, c.d_fluid, 35 * N * sizeof(double),
                              hipMemcpyDefault, c.stream), "pull fluidvars");
  if (ok && fluidauxvars)
    ok = hipok(hipMemcpyAsync(fluidauxvars, c.d_aux, 14 * N * sizeof(double),
                              hipMemcpyDefault, c.stream), "pull fluidauxvars");
  if (ok) ok = hipok(hipStreamSynchronize(c.stream), "pull");
  return ok ? 0 : 2;
}

int gemhip_get_interface_vels(double* vs1i, double* vs2i, double* vs3i) {
  if (!ready()) return 1;
  Ctx& c = ctx();
  if (!c.vsi_current) {
    fail("interface velocities were formed inside gemhip_transport_step and not materialised");
    return 2;
  }
  bool ok = true;
  if (vs1i) ok = ok && hipok(hipMemcpy(vs1i, c.s.vs1i, 7 * nv1i() * sizeof(double),
                                       hipMemcpyDefault), "vs1i");
  if (vs2i) ok = ok && hipok(hipMemcpy(vs2i, c.s.vs2i, 7 * nv2i() * sizeof(double),
                                       hipMemcpyDefault), "vs2i");
  if (vs3i) ok = ok && hipok(hipMemcpy(vs3i, c.s.vs3i, 7 * nv3i() * sizeof(double),
                                       hipMemcpyDefault), "vs3i");
  return ok ? 0 : 3;
}

int gemhip_set_interface_vels(const double* vs1i, const double* vs2i, const double* vs3i) {
  if (!ready()) return 1;
  Ctx& c = ctx();
  if (!vs1i || !vs2i || !vs3i) { fail("null interface velocity pointer"); return 2; }
  hipStreamSynchronize(c.stream);
  bool ok = hipok(hipMemcpy(c.s.vs1i, vs1i, 7 * nv1i() * sizeof(double), hipMemcpyDefault),
                  "vs1i") &&
            hipok(hipMemcpy(c.s.vs2i, vs2i, 7 * nv2i() * sizeof(double), hipMemcpyDefault),
                  "vs2i") &&
            hipok(hipMemcpy(c.s.vs3i, vs3i, 7 * nv3i() * sizeof(double), hipMemcpyDefault),
                  "vs3i");
  if (!ok) return 3;
  // x1 boundary planes v1i(1,:,:,:), v1i(lx1+1,:,:,:) for the ghost fill
  const long lines = (long)c.lx2 * c.lx3;
  std::vector<double> b(14 * lines);
  for (int isp = 0; isp < 7; ++isp)
    for (long l = 0; l < lines; ++l) {
      const double* v = vs1i + isp * nv1i() + l * (c.lx1 + 1);
      b[2 * (l + lines * isp)] = v[0];
      b[2 * (l + lines * isp) + 1] = v[c.lx1];
    }
  if (!hipok(hipMemcpy(c.s.v1b, b.data(), b.size() * sizeof(double), hipMemcpyDefault),
             "v1b"))
    return 4;
  c.vsi_current = true;
  return 0;
}

}  // extern "C"

namespace gemhip {
namespace {
// tape of one recorded step (doubles): NE0, ES3, ES1, ES2, NE1, TD (ghosted electron
// fields), then the electron interface velocities V1, V2, V3
// the recorded fields, then the electron columns' pivot record (ints, padded to doubles)
long tape_doubles() {
  return 6L * ctx().g.N + nv1i() + nv2i() + nv3i() + (pivot_record_ints(ctx().g) + 1) / 2;
}

AdjTape tape_view(const double* t) {
  const long N = ctx().g.N;
  AdjTape v;
  v.NE0 = t; v.ES3 = t + N; v.ES1 = t + 2 * N; v.ES2 = t + 3 * N; v.NE1 = t + 4 * N;
  v.TD = t + 5 * N; v.V1 = t + 6 * N; v.V2 = v.V1 + nv1i(); v.V3 = v.V2 + nv2i();
  v.PIV = reinterpret_cast<int*>(const_cast<double*>(v.V3 + nv3i()));
  return v;
}

// One transport step on the HBM state.  With a tape, the electron fields the adjoint needs
// are copied out on the way (device-to-device, on the library stream); the step itself is
// the same sequence of kernels either way.
int run_step(double dt, double* tape) {
  if (!ready(true)) return 1;
  Ctx& c = ctx();
  if (c.residency != 1) { fail("gemhip_transport_step needs residency mode 1"); return 2; }
  const DevGrid& g = c.g;
  const DevState& s = c.s;
  hipStream_t st = c.stream;
  const long eo = (long)(LSP - 1) * g.N;             // electron field offset
  const size_t fb = (size_t)g.N * sizeof(double);
  auto keep = [&](long slot, const double* src) {
    hipMemcpyAsync(tape + slot * g.N, src + eo, fb, hipMemcpyDeviceToDevice, st);
  };
  // libgemini.cpp:155-156, 168-176: with all sides global and no periodic wrap, the
  // conversions fold into the ghost fill and the x3 sweep (the adjoint's tape needs the
  // converted energy entering the x3 sweep, so a recorded step converts as a pass)
  const bool fuse = !tape && preconvert_fusable(g);
  if (fuse) {
    launch_set_global_boundaries_preconvert(g, s, st);
    launch_save_v1_boundaries(g, s, st);
    launch_sweep3_preconvert(g, s, dt, st, c.tune);
  } else {
    launch_v12rhov1_T2rhoe(g, s, true, true, st);        // libgemini.cpp:155-156
    if (tape) keep(0, s.ns);
    launch_set_global_boundaries(g, s, LSP, st);         // :168
    halo_fields({NS, RHOVS1, RHOES, VS2, VS3});          // :169
    launch_save_v1_boundaries(g, s, st);                 // :170 (x1 planes kept for next step)
    if (tape) {
      keep(1, s.rhoes);
      const AdjTape v = tape_view(tape);
      launch_tape_vels(g, s, const_cast<double*>(v.V1), const_cast<double*>(v.V2),
                       const_cast<double*>(v.V3), st);
    }
    launch_sweep(g, s, 3, F_ALL, dt, true, st, c.tune);  // :174-176
  }
  if (tape) keep(2, s.rhoes);
  launch_sweep(g, s, 1, F_ALL, dt, true, st, c.tune);   // :178-180
  halo_fields({NS, RHOVS1, RHOES});                      // :181
  if (tape) keep(3, s.rhoes);
  if (!tape && postconvert_fusable(g)) {                 // :183-186 with the conversions
    launch_sweep2_postconvert(g, s, dt, st, c.tune);     // folded into the x2 sweep
  } else {
    launch_sweep(g, s, 2, F_ALL, dt, true, st, c.tune);  // :183-185
    launch_post_convert(g, s, st);                       // :186, then multifluid.f90:179-189
  }
  if (tape) { keep(4, s.ns); keep(5, s.Ts); }
  if (!hipok(hipGetLastError(), "advection launch")) return 3;
  c.vsi_current = false;
  // a recorded step also keeps the electron columns' row interchanges (written by the
  // fitting extension's kernel, the one a recorded run uses)
  if (tape) c.s.pivrec = tape_view(tape).PIV;
  // :208; the conversions after the x2 sweep floored every ghost temperature at 100 K
  // already (multifluid.f90:340's floor is idempotent there)
  const bool ok = run_energy_diffusion(dt, false);
  c.s.pivrec = nullptr;
  return ok ? 0 : 4;
}
// The transport block of src/libgemini.cpp:155-210: the advection part of run_step
// (the energy density stays as the sweeps leave it), rhov12v1, clean_param 1 and 2,
// VNRicht_artvisc, halo_fluidvars, RK2_global_boundary_allspec, compression, rhoe2T,
// clean_param 3, energy_diffusion, clean_param 3, T2rhoe.
int run_block(double dt) {
  if (!ready(true)) return 1;
  Ctx& c = ctx();
  if (c.residency != 1) { fail("gemhip_fluid_adv_block needs residency mode 1"); return 2; }
  const DevGrid& g = c.g;
  const DevState& s = c.s;
  hipStream_t st = c.stream;
  if (preconvert_fusable(g)) {                           // as in run_step
    launch_set_global_boundaries_preconvert(g, s, st);
    launch_save_v1_boundaries(g, s, st);
    launch_sweep3_preconvert(g, s, dt, st, c.tune);
  } else {
    launch_v12rhov1_T2rhoe(g, s, true, true, st);        // libgemini.cpp:155-156
    launch_set_global_boundaries(g, s, LSP, st);         // :168
    halo_fields({NS, RHOVS1, RHOES, VS2, VS3});          // :169
    launch_save_v1_boundaries(g, s, st);                 // :170
    launch_sweep(g, s, 3, F_ALL, dt, true, st, c.tune);  // :174-176
  }
  launch_sweep(g, s, 1, F_ALL, dt, true, st, c.tune);   // :178-180
  halo_fields({NS, RHOVS1, RHOES});                      // :181
  launch_sweep(g, s, 2, F_ALL, dt, true, st, c.tune);   // :183-185
  const int* nul = c.d_null;
  launch_rhov12v1_clean_density(g, s, nul, st);          // :186-187, one pass
  launch_clean_param(g, s, 2, nul, st);                  // :188
  launch_artvisc(g, s, c.d_Q, st);                       // :191
  halo_fields({NS, RHOVS1, RHOES, VS2, VS3});            // :199
  launch_rk2_global_boundary(g, s, st);                  // :200
  launch_compression_rhoe2T_clean(g, s, c.d_Q, dt, nul, st);   // :203-205, one pass
  if (!hipok(hipGetLastError(), "transport block launch")) return 3;
  c.vsi_current = false;
  if (!run_energy_diffusion(dt)) return 4;               // :208
  launch_clean_T2rhoe(g, s, nul, st);                    // :209-210, one pass
  return hipok(hipGetLastError(), "transport block launch") ? 0 : 3;
}

// d_src layout (ni = lx1*lx2*lx3): nn 6 ni | vn1 vn2 vn3 | Prionize 6 ni | Qeionize | g1 |
// alt (N, ghosted)
constexpr long SRC_NI = 18;
bool src_alloc() {
  Ctx& c = ctx();
  if (c.d_src) return true;
  const long ni = (long)c.lx1 * c.lx2 * c.lx3;
  return dev_alloc(&c.d_src, (size_t)(SRC_NI * ni + c.g.N), "source/loss inputs");
}

DevSources sources() {
  Ctx& c = ctx();
  const long ni = (long)c.lx1 * c.lx2 * c.lx3;
  DevSources a;
  a.nn = c.d_src;
  a.vn1 = c.d_src + 6 * ni;
  a.vn2 = c.d_src + 7 * ni;
  a.vn3 = c.d_src + 8 * ni;
  a.Prionize = c.d_src + 9 * ni;
  a.Qeionize = c.d_src + 15 * ni;
  a.g1 = c.d_src + 16 * ni;
  a.alt = c.d_src + SRC_NI * ni;
  a.E1 = c.d_J1 + c.g.N;
  a.Q = c.d_Q;
  return a;
}

// which: 1 energy, 2 momentum, 4 mass
bool sources_ready(int which) {
  Ctx& c = ctx();
  if (!c.atmos_ok) { fail("neutral atmosphere not registered (gemhip_register_atmosphere)"); return false; }
  if ((which & 2) && !c.gravity_ok) {
    fail("altitude and gravity not registered (gemhip_register_gravity)");
    return false;
  }
  if ((which & 2) && c.lx1 < 2) { fail("the momentum sources need lx1 >= 2"); return false; }
  return true;
}

void launch_sources(int which, double dt) {
  Ctx& c = ctx();
  const DevSources a = sources();
  if (which & 1) launch_source_loss_energy(c.g, c.s, a, dt, c.stream);
  if (which & 2) launch_source_loss_momentum(c.g, c.s, a, dt, c.stream);
  if (which & 4) launch_source_loss_mass(c.g, c.s, a, dt, c.stream);
}

// src/libgemini.cpp:223-232 after the ionization rates are known
int run_sources_block(double dt) {
  if (!ready(true) || !sources_ready(7)) return 1;
  Ctx& c = ctx();
  if (c.residency != 1) { fail("gemhip_sources_block needs residency mode 1"); return 2; }
  const DevGrid& g = c.g;
  const DevState& s = c.s;
  hipStream_t st = c.stream;
  const int* nul = c.d_null;
  launch_set_global_boundaries(g, s, LSP, st);           // :223
  launch_sources(3, dt);                                 // :225-226
  // :227 with clean_param 3 and 1 of :229-231 in the same pass (they touch Ts and ns;
  // clean_param 2 between them touches vs1 only)
  launch_source_loss_mass(g, s, sources(), dt, st, nul, true);
  launch_clean_param(g, s, 2, nul, st);
  return hipok(hipGetLastError(), "sources block launch") ? 0 : 3;
}

// one source_loss_*_C call: 1 energy, 2 momentum, 4 mass, 7 all three in that order
void source_entry(double** fv, double** av, double** ev, int* xtype, double* dt, int which) {
  if (!ready(true) || !check_xtype(xtype) || !check_host(fv, "fluidvars") ||
      !check_host(av, "fluidauxvars") || !sources_ready(which))
    return;
  if (!dt) { fail("dt is a null pointer"); return; }
  if ((which & 2) && !check_host(ev, "electrovars")) return;
  Io io{fv, av, ev};
  bool ok = upload(io, {NS, VS1, VS2, VS3, TS});
  if (ok && (which & 1)) ok = upload(io, {RHOES});
  if (ok && (which & 2)) ok = upload(io, {RHOVS1, E1});
  if (!ok) return;
  launch_sources(which, *dt);
  if (!hipok(hipGetLastError(), "source/loss launch")) return;
  if (which & 1) download(io, {TS, RHOES});
  if (which & 2) download(io, {VS1, RHOVS1});
  if (which & 4) download(io, {NS});
}
}  // namespace
}  // namespace gemhip

extern "C" {

int gemhip_transport_step(double dt) { return run_step(dt, nullptr); }

int gemhip_fluid_adv_block(double dt) { return run_block(dt); }

int gemhip_sources_block(double dt) { return run_sources_block(dt); }

int gemhip_register_atmosphere(const double* nn, const double* Tn, const double* vn1,
                               const double* vn2, const double* vn3) {
  Ctx& c = ctx();
  if (!ready()) return 1;
  if (!nn || !Tn || !vn1 || !vn2 || !vn3) { fail("null atmosphere pointer"); return 2; }
  if (!src_alloc()) return 3;
  const long ni = (long)c.lx1 * c.lx2 * c.lx3;
  const size_t b = (size_t)ni * sizeof(double);
  const DevSources a = sources();
  bool ok = hipok(hipMemcpy(c.d_src, nn, 6 * b, hipMemcpyDefault), "nn upload") &&
            hipok(hipMemcpy(const_cast<double*>(a.vn1), vn1, b, hipMemcpyDefault), "vn1 upload") &&
            hipok(hipMemcpy(const_cast<double*>(a.vn2), vn2, b, hipMemcpyDefault), "vn2 upload") &&
            hipok(hipMemcpy(const_cast<double*>(a.vn3), vn3, b, hipMemcpyDefault), "vn3 upload") &&
            hipok(hipMemcpy(c.d_neut, nn, 3 * b, hipMemcpyDefault), "nn upload") &&
            hipok(hipMemcpy(c.d_neut + 3 * ni, Tn, b, hipMemcpyDefault), "Tn upload");
  if (!ok) return 4;
  c.atmos_ok = c.neutrals_ok = true;
  return 0;
}

int gemhip_register_gravity(const double* alt, const double* g1) {
  Ctx& c = ctx();
  if (!ready()) return 1;
  if (!alt || !g1) { fail("null altitude or gravity pointer"); return 2; }
  if (!src_alloc()) return 3;
  const long ni = (long)c.lx1 * c.lx2 * c.lx3;
  const DevSources a = sources();
  bool ok = hipok(hipMemcpy(const_cast<double*>(a.alt), alt, (size_t)c.g.N * sizeof(double),
                            hipMemcpyDefault), "alt upload") &&
            hipok(hipMemcpy(const_cast<double*>(a.g1), g1, (size_t)ni * sizeof(double),
                            hipMemcpyDefault), "g1 upload");
  if (!ok) return 4;
  c.gravity_ok = true;
  return 0;
}

int gemhip_set_ionization_rates(const double* Prionize, const double* Qeionize) {
  Ctx& c = ctx();
  if (!ready()) return 1;
  if (!Prionize || !Qeionize) { fail("null ionization-rate pointer"); return 2; }
  if (!src_alloc()) return 3;
  const long ni = (long)c.lx1 * c.lx2 * c.lx3;
  const DevSources a = sources();
  bool ok = hipok(hipMemcpyAsync(const_cast<double*>(a.Prionize), Prionize,
                                 6 * (size_t)ni * sizeof(double), hipMemcpyDefault, c.stream),
                  "Prionize upload") &&
            hipok(hipMemcpyAsync(const_cast<double*>(a.Qeionize), Qeionize,
                                 (size_t)ni * sizeof(double), hipMemcpyDefault, c.stream),
                  "Qeionize upload") &&
            hipok(hipStreamSynchronize(c.stream), "ionization rates");
  return ok ? 0 : 4;
}

void clear_ionization_arrays_C(void** intvars) {
  (void)intvars;
  if (!ready() || !src_alloc()) return;
  Ctx& c = ctx();
  const long ni = (long)c.lx1 * c.lx2 * c.lx3;
  launch_zero(const_cast<double*>(sources().Prionize), 7 * ni, c.stream);   // Prionize, Qeionize
  hipok(hipGetLastError(), "clear ionization arrays");
  if (c.residency == 0) hipok(hipStreamSynchronize(c.stream), "clear ionization arrays");
}

void source_loss_energy_C(double** fluidvars, double** fluidauxvars, double** electrovars,
                          void** intvars, int* xtype, void** xC, double* dt) {
  (void)intvars; (void)xC;
  source_entry(fluidvars, fluidauxvars, electrovars, xtype, dt, 1);
}

void source_loss_momentum_C(double** fluidvars, double** fluidauxvars, double** electrovars,
                            void** intvars, int* xtype, void** xC, double* dt) {
  (void)intvars; (void)xC;
  source_entry(fluidvars, fluidauxvars, electrovars, xtype, dt, 2);
}

void source_loss_mass_C(double** fluidvars, double** fluidauxvars, double** electrovars,
                        void** intvars, int* xtype, void** xC, double* dt) {
  (void)intvars; (void)xC;
  source_entry(fluidvars, fluidauxvars, electrovars, xtype, dt, 4);
}

void source_loss_allparams_C(double** fluidvars, double** fluidauxvars, double** electrovars,
                             void** intvars, int* xtype, void** xC, double* dt) {
  (void)intvars; (void)xC;
  source_entry(fluidvars, fluidauxvars, electrovars, xtype, dt, 7);
}

int gemhip_register_cfl(double tcfl, int flagglow) {
  Ctx& c = ctx();
  c.tcfl = tcfl;
  c.flagglow = flagglow;
  c.dt_calls = 0;
  return 0;
}

int gemhip_grid_rates(double* rates) {
  if (!ready()) return 1;
  Ctx& c = ctx();
  if (!rates) { fail("null rates pointer"); return 2; }
  const int nb = gridrate_blocks(c.g);
  if (!c.d_rates && !hipok(hipMalloc((void**)&c.d_rates, (size_t)(nb + 1) * 3 * LSP * sizeof(double)),
                           "grid rates"))
    return 3;
  launch_gridrates(c.g, c.s, c.d_rates, c.stream);
  if (!hipok(hipMemcpyAsync(rates, c.d_rates + (long)nb * 3 * LSP, 3 * LSP * sizeof(double),
                            hipMemcpyDefault, c.stream), "grid rates") ||
      !hipok(hipStreamSynchronize(c.stream), "grid rates"))
    return 4;
  return 0;
}

// libgemini_mpi_c.f90:289-308 -> libgemini_mpi.f90:450-483, temporal_mpi.f90:27-165 on one
// process.  The reference's iteration count (get_it) is the number of dt_select calls.
void dt_select_C(void** cfgC, int* xtype, void** xC, double** fluidvars, double** fluidauxvars,
                 double* t, double* tout, double* tglowout, double* dt) {
  (void)cfgC; (void)xC; (void)fluidauxvars;
  if (!ready() || !check_xtype(xtype) || !check_host(fluidvars, "fluidvars")) return;
  if (!t || !tout || !tglowout || !dt) { fail("null time argument"); return; }
  Ctx& c = ctx();
  Io io{fluidvars, nullptr, nullptr};
  if (!upload(io, {TS, VS1, VS2, VS3})) return;
  double r[3 * LSP];
  if (gemhip_grid_rates(r) != 0) return;
  const double dtprev = *dt;
  double m = 0.0;
  for (int k = 0; k < 3; ++k) {                  // min(minval(1/gridrate1), ..2, ..3)
    double mk = 0.0;
    for (int isp = 1; isp <= LSP; ++isp) {
      const double inv = 1.0 / fmax_f(r[3 * (isp - 1) + k], 1e-10);
      if (isp == 1 || inv < mk) mk = inv;
    }
    if (k == 0 || mk < m) m = mk;
  }
  double v = c.tcfl * m;
  if (c.flagglow != 0 && *t + v > *tglowout) v = *tglowout - *t;
  if (*t + v > *tout) v = *tout - *t;
  v = fmax_f(v, 1e-6);
  c.dt_calls += 1;
  if (c.dt_calls > 1 && v / dtprev > 2.0) v = 2.0 * dtprev;
  *dt = v;
}

int gemhip_register_null_cells(const int* nullpts) {
  if (!ready()) return 1;
  Ctx& c = ctx();
  hipStreamSynchronize(c.stream);
  if (!nullpts) {
    if (c.d_null) hipFree(c.d_null);
    c.d_null = nullptr;
    return 0;
  }
  const size_t n = (size_t)c.lx1 * c.lx2 * c.lx3;
  std::vector<int> h(n);                          // a mask without null cells is no mask
  if (!hipok(hipMemcpy(h.data(), nullpts, n * sizeof(int), hipMemcpyDefault), "null cells"))
    return 3;
  size_t count = 0;
  for (int v : h) count += v != 0;
  if (count == 0) return gemhip_register_null_cells(nullptr);
  if (!c.d_null && !hipok(hipMalloc((void**)&c.d_null, n * sizeof(int)), "null cells")) return 2;
  if (!hipok(hipMemcpy(c.d_null, h.data(), n * sizeof(int), hipMemcpyHostToDevice),
             "null cells upload"))
    return 3;
  return 0;
}

void clean_param_C(int* iparm, int* xtype, void** xC, double** fluidvars) {
  (void)xC;
  if (!ready() || !check_xtype(xtype) || !check_host(fluidvars, "fluidvars")) return;
  if (!iparm || *iparm < 1 || *iparm > 3) {     // libgemini.f90:1050
    fail("  libgemini:clean_params_C(); invalid parameter selected");
    return;
  }
  Ctx& c = ctx();
  const Var v = *iparm == 1 ? NS : *iparm == 2 ? VS1 : TS;
  Io io{fluidvars, nullptr, nullptr};
  if (!upload(io, {v})) return;
  launch_clean_param(c.g, c.s, *iparm, c.d_null, c.stream);
  if (!hipok(hipGetLastError(), "clean_param launch")) return;
  download(io, {v});
}

void VNRicht_artvisc_C(double** fluidvars, void** intvars) {
  (void)intvars;
  if (!ready() || !check_host(fluidvars, "fluidvars")) return;
  Ctx& c = ctx();
  Io io{fluidvars, nullptr, nullptr};
  if (!upload(io, {NS, VS1})) return;
  launch_artvisc(c.g, c.s, c.d_Q, c.stream);
  if (c.residency == 0) hipok(hipStreamSynchronize(c.stream), "artificial viscosity");
}

void compression_C(double** fluidvars, double** fluidauxvars, void** intvars, int* xtype,
                   void** xC, double* dt) {
  (void)intvars; (void)xC;
  if (!ready() || !check_xtype(xtype) || !check_host(fluidvars, "fluidvars") ||
      !check_host(fluidauxvars, "fluidauxvars"))
    return;
  if (!dt) { fail("dt is a null pointer"); return; }
  Ctx& c = ctx();
  Io io{fluidvars, fluidauxvars, nullptr};
  if (!upload(io, {VS1, VS2, VS3, RHOES})) return;
  launch_compression(c.g, c.s, c.d_Q, *dt, c.stream);
  if (!hipok(hipGetLastError(), "compression launch")) return;
  download(io, {RHOES});
}

void RK2_prep_mpi_allspec_C(int* xtype, void** xC, double** fluidvars) {
  (void)xC;
  if (!ready() || !check_xtype(xtype)) return;
  Ctx& c = ctx();
  if (!c.g.flagper) return;                    // one process, aperiodic: nothing to exchange
  if (!check_host(fluidvars, "fluidvars")) return;
  Io io{fluidvars, nullptr, nullptr};
  if (!upload(io, {VS1, VS2, VS3})) return;
  double* f[3] = {dev_var(VS1), dev_var(VS2), dev_var(VS3)};
  launch_halo_periodic(c.g, f, 3, 1, LSP, c.stream);   // sources_mpi.f90:14-50, depth 1
  download(io, {VS1, VS2, VS3});
}

void RK2_global_boundary_allspec_C(int* xtype, void** xC, double** fluidvars) {
  (void)xC;
  if (!ready() || !check_xtype(xtype) || !check_host(fluidvars, "fluidvars")) return;
  Ctx& c = ctx();
  Io io{fluidvars, nullptr, nullptr};
  if (!upload(io, {VS1, VS2, VS3})) return;
  launch_rk2_global_boundary(c.g, c.s, c.stream);
  if (!hipok(hipGetLastError(), "RK2 boundary launch")) return;
  download(io, {VS1, VS2, VS3});
}

// ---- part 4: parameter fitting ----
int gemhip_register_conductivity_scale(const double* kappa) {
  if (!ready()) return 1;
  Ctx& c = ctx();
  hipStreamSynchronize(c.stream);
  if (!kappa) { c.s.kappa = nullptr; return 0; }
  const size_t n = (size_t)c.lx2 * c.lx3;
  if (!c.d_kappa && !hipok(hipMalloc((void**)&c.d_kappa, n * sizeof(double)), "kappa")) return 2;
  if (!hipok(hipMemcpy(c.d_kappa, kappa, n * sizeof(double), hipMemcpyDefault), "kappa upload"))
    return 3;
  c.s.kappa = c.d_kappa;
  return 0;
}

int gemhip_register_top_heat_flux(const double* qtop) {
  if (!ready()) return 1;
  Ctx& c = ctx();
  hipStreamSynchronize(c.stream);
  if (!qtop) { c.s.qtop = nullptr; return 0; }
  if (c.g.gridflag != 2) {
    fail("a top heat flux needs gridflag 2 (top at x1 = lx1), got %d", c.g.gridflag);
    return 2;
  }
  const size_t n = (size_t)c.lx2 * c.lx3;
  if (!c.d_qtop && !hipok(hipMalloc((void**)&c.d_qtop, n * sizeof(double)), "qtop")) return 3;
  if (!hipok(hipMemcpy(c.d_qtop, qtop, n * sizeof(double), hipMemcpyDefault), "qtop upload"))
    return 4;
  c.s.qtop = c.d_qtop;
  return 0;
}

long gemhip_checkpoint_size(void) {
  Ctx& c = ctx();
  return c.sized ? 35L * c.g.N + 14L * c.lx2 * c.lx3 : 0;
}

int gemhip_checkpoint_save(double* dst) {
  if (!ready()) return 1;
  Ctx& c = ctx();
  if (!dst) { fail("null checkpoint pointer"); return 2; }
  const size_t nf = 35 * (size_t)c.g.N, nb = 14 * (size_t)c.lx2 * c.lx3;
  bool ok = hipok(hipMemcpyAsync(dst, c.d_fluid, nf * sizeof(double), hipMemcpyDefault,
                                 c.stream), "checkpoint fluidvars") &&
            hipok(hipMemcpyAsync(dst + nf, c.d_v1b, nb * sizeof(double), hipMemcpyDefault,
                                 c.stream), "checkpoint v1b") &&
            hipok(hipStreamSynchronize(c.stream), "checkpoint");
  return ok ? 0 : 3;
}

int gemhip_checkpoint_load(const double* src) {
  if (!ready()) return 1;
  Ctx& c = ctx();
  if (!src) { fail("null checkpoint pointer"); return 2; }
  const size_t nf = 35 * (size_t)c.g.N, nb = 14 * (size_t)c.lx2 * c.lx3;
  bool ok = hipok(hipMemcpyAsync(c.d_fluid, src, nf * sizeof(double), hipMemcpyDefault,
                                 c.stream), "restore fluidvars") &&
            hipok(hipMemcpyAsync(c.d_v1b, src + nf, nb * sizeof(double), hipMemcpyDefault,
                                 c.stream), "restore v1b") &&
            hipok(hipStreamSynchronize(c.stream), "restore");
  c.vsi_current = false;
  return ok ? 0 : 3;
}

long gemhip_adjoint_tape_size(void) { return ctx().sized ? tape_doubles() : 0; }

int gemhip_adjoint_record_step(double dt, double* tape) {
  if (!tape) { fail("null tape pointer"); return 5; }
  return run_step(dt, tape);
}

int gemhip_adjoint_step(double dt, const double* tape, double* lam, double* gkappa,
                        double* gq) {
  if (!ready(true)) return 1;
  Ctx& c = ctx();
  if (!tape || !lam) { fail("null tape or adjoint pointer"); return 2; }
  if (c.diffsolvetype != 2) { fail("the adjoint needs the TRBDF2 solver (diffsolvetype 2)"); return 3; }
  if (c.g.flagper && c.lx3 < 2) { fail("the adjoint of the periodic wrap needs lx3 >= 2"); return 4; }
  const size_t need = (size_t)c.g.N * sizeof(double) + adjoint_scratch_bytes(c.g);
  if (!c.d_adj || c.adj_bytes < need) {
    hipStreamSynchronize(c.stream);
    if (c.d_adj) hipFree(c.d_adj);
    c.d_adj = nullptr;
    c.adj_bytes = 0;
    if (!hipok(hipMalloc((void**)&c.d_adj, need), "adjoint work")) return 5;
    c.adj_bytes = need;
  }
  launch_adjoint_step(c.g, c.s, tape_view(tape), dt, c.Teinf, lam, c.d_adj, gkappa,
                      c.s.qtop ? gq : nullptr, c.d_adj + c.g.N, c.stream);
  return hipok(hipGetLastError(), "adjoint launch") ? 0 : 6;
}

int gemhip_device_state(double** fluidvars, double** rhovs1_rhoes, double** J1p) {
  if (!ready()) return 1;
  Ctx& c = ctx();
  if (fluidvars) *fluidvars = c.d_fluid;
  if (rhovs1_rhoes) *rhovs1_rhoes = c.d_aux;
  if (J1p) *J1p = c.d_J1;
  return 0;
}

void* gemhip_stream(void) { return (void*)ctx().stream; }

int gemhip_synchronize(void) {
  Ctx& c = ctx();
  if (!c.inited) return 1;
  return hipok(hipStreamSynchronize(c.stream), "synchronise") ? 0 : 2;
}

int gemhip_kernel_timing(int enable) {
  Ctx& c = ctx();
  if (!enable && c.inited) { hipStreamSynchronize(c.stream); collect_timings(); }
  c.timing = enable != 0;
  return 0;
}

int gemhip_kernel_time(const char* name, double* total_ms, long* launches) {
  Ctx& c = ctx();
  if (c.inited) collect_timings();
  auto it = c.acc.find(name ? name : "");
  if (it == c.acc.end()) {
    if (total_ms) *total_ms = 0.0;
    if (launches) *launches = 0;
    return 1;
  }
  if (total_ms) *total_ms = it->second.first;
  if (launches) *launches = it->second.second;
  return 0;
}

int gemhip_kernel_timing_reset(void) {
  Ctx& c = ctx();
  if (c.inited) { hipStreamSynchronize(c.stream); collect_timings(); }
  c.acc.clear();
  return 0;
}

// ------------------------------------------------------------ reference-named entries

void v12rhov1_C(double** fluidvars, double** fluidauxvars) {
  if (!ready() || !check_host(fluidvars, "fluidvars") || !check_host(fluidauxvars, "fluidauxvars"))
    return;
  Ctx& c = ctx();
  Io io{fluidvars, fluidauxvars, nullptr};
  if (!upload(io, {NS, VS1})) return;
  launch_v12rhov1_T2rhoe(c.g, c.s, true, false, c.stream);
  download(io, {RHOVS1});
}

void T2rhoe_C(double** fluidvars, double** fluidauxvars) {
  if (!ready() || !check_host(fluidvars, "fluidvars") || !check_host(fluidauxvars, "fluidauxvars"))
    return;
  Ctx& c = ctx();
  Io io{fluidvars, fluidauxvars, nullptr};
  if (!upload(io, {NS, TS})) return;
  launch_v12rhov1_T2rhoe(c.g, c.s, false, true, c.stream);
  download(io, {RHOES});
}

void rhoe2T_C(double** fluidvars, double** fluidauxvars) {
  if (!ready() || !check_host(fluidvars, "fluidvars") || !check_host(fluidauxvars, "fluidauxvars"))
    return;
  Ctx& c = ctx();
  Io io{fluidvars, fluidauxvars, nullptr};
  if (!upload(io, {NS, RHOES})) return;
  launch_rhoe2T(c.g, c.s, c.stream);
  download(io, {TS});
}

void rhov12v1_C(double** fluidvars, double** fluidauxvars) {
  if (!ready() || !check_host(fluidvars, "fluidvars") || !check_host(fluidauxvars, "fluidauxvars"))
    return;
  Ctx& c = ctx();
  Io io{fluidvars, fluidauxvars, nullptr};
  if (!upload(io, {NS, RHOVS1})) return;
  launch_rhov12v1(c.g, c.s, c.stream);
  download(io, {NS, VS1});
}

void interface_vels_allspec_C(int* xtype, void** xC, double** fluidvars, void** intvars, int* lsp) {
  (void)xC; (void)intvars;
  if (!ready() || !check_xtype(xtype) || !check_host(fluidvars, "fluidvars")) return;
  if (!lsp || *lsp > 7 || *lsp < 0) {
    fail("number of interface vels must be less than or equal to total species number");
    return;
  }
  Ctx& c = ctx();
  Io io{fluidvars, nullptr, nullptr};
  if (!upload(io, {VS1, VS2, VS3})) return;
  if (*lsp > 0) launch_interface_vels(c.g, c.s, *lsp, c.stream);
  c.vsi_current = true;
  if (c.residency == 0) hipok(hipStreamSynchronize(c.stream), "synchronise");
}

void set_global_boundaries_allspec_C(int* xtype, void** xC, double** fluidvars,
                                     double** fluidauxvars, void** intvars, int* lsp) {
  (void)xC; (void)intvars;
  if (!ready() || !check_xtype(xtype) || !check_host(fluidvars, "fluidvars") ||
      !check_host(fluidauxvars, "fluidauxvars"))
    return;
  if (!lsp || *lsp > 7 || *lsp < 0) {
    fail("number of global boundaries must be less than or equal to total species number");
    return;
  }
  Ctx& c = ctx();
  Io io{fluidvars, fluidauxvars, nullptr};
  if (!upload(io, {NS, VS1, VS2, VS3, RHOVS1, RHOES})) return;
  if (*lsp > 0) launch_set_global_boundaries(c.g, c.s, *lsp, c.stream);
  download(io, {NS, VS2, VS3, RHOVS1, RHOES});
}

#define SWEEP_ENTRY(NAME, DIR, FIELDS)                                                        \
  void NAME(double** fv, double** av, void** intvars, int* xtype, void** xC, double* dt) {   \
    (void)intvars; (void)xC;                                                                  \
    sweep_entry(fv, av, xtype, dt, DIR, FIELDS);                                              \
  }
SWEEP_ENTRY(sweep3_allparams_C, 3, F_ALL)
SWEEP_ENTRY(sweep3_allspec_mass_C, 3, F_MASS)
SWEEP_ENTRY(sweep3_allspec_momentum_C, 3, F_MOMENTUM)
SWEEP_ENTRY(sweep3_allspec_energy_C, 3, F_ENERGY)
SWEEP_ENTRY(sweep1_allparams_C, 1, F_ALL)
SWEEP_ENTRY(sweep1_allspec_mass_C, 1, F_MASS)
SWEEP_ENTRY(sweep1_allspec_momentum_C, 1, F_MOMENTUM)
SWEEP_ENTRY(sweep1_allspec_energy_C, 1, F_ENERGY)
SWEEP_ENTRY(sweep2_allparams_C, 2, F_ALL)
SWEEP_ENTRY(sweep2_allspec_mass_C, 2, F_MASS)
SWEEP_ENTRY(sweep2_allspec_momentum_C, 2, F_MOMENTUM)
SWEEP_ENTRY(sweep2_allspec_energy_C, 2, F_ENERGY)
#undef SWEEP_ENTRY

void energy_diffusion_C(void** cfgC, int* xtype, void** xC, double** fluidvars,
                        double** electrovars, void** intvars, double* dt) {
  (void)cfgC; (void)xC; (void)intvars;
  if (!ready(true) || !check_xtype(xtype) || !check_host(fluidvars, "fluidvars") ||
      !check_host(electrovars, "electrovars"))
    return;
  if (!dt) { fail("dt is a null pointer"); return; }
  Ctx& c = ctx();
  Io io{fluidvars, nullptr, electrovars};
  if (!upload(io, {NS, TS, J1})) return;
  if (!run_energy_diffusion(*dt)) return;
  download(io, {TS});
}

// energy_diffusion_C's kernels on device fields the caller owns, with the work space in the
// static `scratch` array instead of library allocations: Ts and ns are the 7 ghosted species
// fields of fluidvars(:,:,:,29:35) and (:,:,:,1:7), J1 one ghosted field, all in HBM.  The
// selected context supplies the grid, neutrals and configuration.  Enqueued on the
// context's stream (gemhip_synchronize waits for it).  Every pivot row, right-hand side,
// solution and interchange word is written before it is read within a call, so the work
// space needs no clearing; calls that overlap on one GPU share it and race.
int gemhip_energy_diffusion_noalloc(double* Ts, const double* ns, const double* J1, double dt) {
  if (!ready(true)) return 1;
  Ctx& c = ctx();
  if (!Ts || !ns || !J1) { fail("null field pointer"); return 2; }
  if (c.timing) {
    fail("kernel timing creates events on demand: disable it before allocation-free calls");
    return 3;
  }
  const size_t need = diffusion_scratch_bytes(c.g);
  if (need > (size_t)NOALLOC_SCRATCH_BYTES) {
    fail("grid %dx%dx%d needs %zu bytes of diffusion work space; the allocation-free entry "
         "point has %zu", c.lx1, c.lx2, c.lx3, need, (size_t)NOALLOC_SCRATCH_BYTES);
    return 4;
  }
  if (c.device >= MAX_DEVICES || !noalloc_scratch[c.device]) {
    fail("no work space address on device %d: %s", c.device,
         c.device < MAX_DEVICES ? hipGetErrorString(noalloc_lookup[c.device]) : "device ordinal");
    return 5;
  }
  if (c.diffsolvetype != 1 && c.diffsolvetype != 2) {
    fail("Unsupported diffusion solver type/mode:  %d.  Should be either 1 or 2.",
         c.diffsolvetype);
    return 6;
  }
  if (c.lx1 < 4) { fail("energy diffusion needs lx1 >= 4 (got %d)", c.lx1); return 6; }
  if ((c.s.kappa || c.s.qtop) && c.diffsolvetype != 2) {
    fail("the conductivity scale and top heat flux need the TRBDF2 solver (diffsolvetype 2)");
    return 6;
  }
  DevState s = c.s;
  s.Ts = Ts;
  s.ns = const_cast<double*>(ns);
  s.J1 = J1;
  launch_energy_diffusion(c.g, s, dt, c.diffsolvetype, c.Teinf,
                          static_cast<double*>(noalloc_scratch[c.device]), c.stream, c.tune);
  return hipok(hipGetLastError(), "energy diffusion launch") ? 0 : 7;
}

void halo_allparams_C(int* xtype, void** xC, double** fluidvars, double** fluidauxvars) {
  (void)xC;
  if (!ready() || !check_xtype(xtype)) return;
  Ctx& c = ctx();
  if (!c.g.flagper) return;                    // one process, aperiodic: nothing to exchange
  if (!check_host(fluidvars, "fluidvars") || !check_host(fluidauxvars, "fluidauxvars")) return;
  Io io{fluidvars, fluidauxvars, nullptr};
  if (!upload(io, {NS, RHOVS1, RHOES})) return;
  halo_fields({NS, RHOVS1, RHOES});
  download(io, {NS, RHOVS1, RHOES});
}

void halo_fluidvars_C(int* xtype, void** xC, double** fluidvars, double** fluidauxvars) {
  (void)xC;
  if (!ready() || !check_xtype(xtype)) return;
  Ctx& c = ctx();
  if (!c.g.flagper) return;
  if (!check_host(fluidvars, "fluidvars") || !check_host(fluidauxvars, "fluidauxvars")) return;
  Io io{fluidvars, fluidauxvars, nullptr};
  if (!upload(io, {NS, RHOVS1, RHOES, VS2, VS3})) return;
  halo_fields({NS, RHOVS1, RHOES, VS2, VS3});
  download(io, {NS, RHOVS1, RHOES, VS2, VS3});
}

// advec_mpi.f90:14-44: vs2 and vs3 with one ghost layer, species 1..lsp
void halo_interface_vels_allspec_C(int* xtype, void** xC, double** fluidvars, int* lsp) {
  (void)xC;
  if (!ready() || !check_xtype(xtype)) return;
  if (!lsp || *lsp > 7 || *lsp < 0) {
    fail("number of haloed species must be less than or equal to total species number");
    return;
  }
  Ctx& c = ctx();
  if (!c.g.flagper) return;
  if (!check_host(fluidvars, "fluidvars")) return;
  Io io{fluidvars, nullptr, nullptr};
  if (!upload(io, {VS2, VS3})) return;
  double* f[2] = {dev_var(VS2), dev_var(VS3)};
  if (*lsp > 0) launch_halo_periodic(c.g, f, 2, 1, *lsp, c.stream);
  download(io, {VS2, VS3});
}

// ---------------- contexts, tuning and ensembles ----------------

int gemhip_context_create(void) {
  contexts().push_back(new Ctx());
  return (int)contexts().size();
}

int gemhip_context_select(int handle) {
  if (handle == 0) { g_current = nullptr; g_current_handle = 0; return 0; }
  if (handle < 0 || handle > (int)contexts().size() || !contexts()[handle - 1]) {
    fail("no library context %d", handle);
    return 1;
  }
  g_current = contexts()[handle - 1];
  g_current_handle = handle;
  return 0;
}

int gemhip_context_current(void) { return g_current_handle; }

bool context_in_graph(int handle);
bool context_in_batch(int handle);

int gemhip_context_destroy(int handle) {
  if (handle <= 0 || handle > (int)contexts().size() || !contexts()[handle - 1]) return 1;
  // a captured ensemble graph holds this context's HBM and stream: release it first
  if (context_in_graph(handle)) {
    fail("library context %d belongs to a captured ensemble graph (gemhip_ensemble_release)",
         handle);
    return 2;
  }
  if (context_in_batch(handle)) {
    fail("library context %d belongs to a batch (gemhip_batch_release)", handle);
    return 2;
  }
  const int keep = g_current_handle == handle ? 0 : g_current_handle;
  gemhip_context_select(handle);
  gemhip_finalize();
  delete contexts()[handle - 1];
  contexts()[handle - 1] = nullptr;
  gemhip_context_select(keep);
  return 0;
}

int gemhip_set_tuning(const char* knob, int value) {
  Ctx& c = ctx();
  const std::string k = knob ? knob : "";
  if (k == "sweep_pf" && (value == 1 || value == 2)) { c.tune.sweep_pf = value; return 0; }
  if (k == "diff_chunk" && (value == 4 || value == 8)) { c.tune.diff_chunk = value; return 0; }
  fail("unknown tuning knob or value: %s = %d", k.c_str(), value);
  return 1;
}

int gemhip_get_tuning(const char* knob) {
  Ctx& c = ctx();
  const std::string k = knob ? knob : "";
  if (k == "sweep_pf") return c.tune.sweep_pf;
  if (k == "diff_chunk") return c.tune.diff_chunk;
  return -1;
}

namespace {
struct EnsembleGraph {
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  hipStream_t origin = nullptr;
  std::vector<hipEvent_t> events;
  std::vector<int> members;    // context handles whose memory and streams the graph uses
  bool owns_origin = true;     // false: the origin is a batch's stream
};
std::vector<EnsembleGraph*>& graphs() {
  static std::vector<EnsembleGraph*> v;
  return v;
}
}  // namespace

bool context_in_graph(int handle) {
  for (const EnsembleGraph* e : graphs())
    if (e)
      for (int h : e->members)
        if (h == handle) return true;
  return false;
}

namespace {

void release_graph(EnsembleGraph* e) {
  if (e->exec) hipGraphExecDestroy(e->exec);
  if (e->graph) hipGraphDestroy(e->graph);
  for (auto ev : e->events) hipEventDestroy(ev);
  if (e->origin && e->owns_origin) hipStreamDestroy(e->origin);
  delete e;
}
}  // namespace

// Capture nsteps transport steps of every member (each on its own context's stream) into
// one graph: a fork event on an origin stream, each member's kernels on its stream, a join
// back onto the origin.  Members must be in residency mode 1.  Returns a graph id >= 1.
int gemhip_ensemble_capture(int n, const int* handles, const double* dt, int nsteps) {
  const int keep = g_current_handle;
  if (n <= 0 || !handles || !dt || nsteps <= 0) { fail("bad ensemble arguments"); return -1; }
  for (int m = 0; m < n; ++m) {               // everything a step allocates, before capture
    if (gemhip_context_select(handles[m])) return -1;
    Ctx& c = ctx();
    if (!ready(true) || c.residency != 1) {
      fail("ensemble member %d is not set up with its state in HBM (residency mode 1)", m);
      gemhip_context_select(keep);
      return -1;
    }
    if (c.timing) {
      fail("disable kernel timing before capturing an ensemble");
      gemhip_context_select(keep);
      return -1;
    }
    if (!ensure_scratch() || !hipok(hipStreamSynchronize(c.stream), "ensemble member sync")) {
      gemhip_context_select(keep);
      return -1;
    }
  }
  auto* e = new EnsembleGraph();
  e->members.assign(handles, handles + n);
  bool ok = hipok(hipStreamCreateWithFlags(&e->origin, hipStreamNonBlocking), "origin stream");
  e->events.resize(n + 1, nullptr);
  for (int m = 0; ok && m <= n; ++m)
    ok = hipok(hipEventCreateWithFlags(&e->events[m], hipEventDisableTiming), "ensemble event");
  ok = ok && hipok(hipStreamBeginCapture(e->origin, hipStreamCaptureModeGlobal), "begin capture");
  ok = ok && hipok(hipEventRecord(e->events[n], e->origin), "fork");
  for (int m = 0; ok && m < n; ++m) {
    gemhip_context_select(handles[m]);
    Ctx& c = ctx();
    ok = hipok(hipStreamWaitEvent(c.stream, e->events[n], 0), "fork wait");
    for (int k = 0; ok && k < nsteps; ++k) ok = gemhip_transport_step(dt[m]) == 0;
    ok = ok && hipok(hipEventRecord(e->events[m], c.stream), "join");
    ok = ok && hipok(hipStreamWaitEvent(e->origin, e->events[m], 0), "join wait");
  }
  hipGraph_t graph = nullptr;
  const hipError_t endc = hipStreamEndCapture(e->origin, &graph);
  e->graph = graph;
  ok = ok && hipok(endc, "end capture");
  ok = ok && hipok(hipGraphInstantiate(&e->exec, e->graph, nullptr, nullptr, 0), "instantiate");
  gemhip_context_select(keep);
  if (!ok) { release_graph(e); return -1; }
  graphs().push_back(e);
  return (int)graphs().size();
}

// Launch a captured ensemble graph (asynchronously, then wait when `wait`).
int gemhip_ensemble_launch(int id, int wait) {
  if (id <= 0 || id > (int)graphs().size() || !graphs()[id - 1]) { fail("no ensemble graph %d", id); return 1; }
  EnsembleGraph* e = graphs()[id - 1];
  if (!hipok(hipGraphLaunch(e->exec, e->origin), "graph launch")) return 2;
  if (wait && !hipok(hipStreamSynchronize(e->origin), "graph wait")) return 3;
  return 0;
}

int gemhip_ensemble_release(int id) {
  if (id <= 0 || id > (int)graphs().size() || !graphs()[id - 1]) return 1;
  hipStreamSynchronize(graphs()[id - 1]->origin);
  release_graph(graphs()[id - 1]);
  graphs()[id - 1] = nullptr;
  return 0;
}

}  // extern "C"

namespace {
// An ensemble advanced by batched kernels (one persistent launch per kernel of the fused
// step for all members), with its member tables in HBM and a stream of its own.
struct Batch {
  std::vector<int> handles;
  std::vector<BatchMember> host;
  BatchMember* d_m = nullptr;
  long* d_pre = nullptr;
  BatchTable t{};
  hipStream_t stream = nullptr;
  int device = 0, cus = 0;
  int sweep_pf = 2, diff_chunk = 8, grid = 1;   // kernel variants and persistent grid size
  std::vector<int> graph_ids;  // graphs captured from this batch (launched on its stream)
};
std::vector<Batch*>& batches() {
  static std::vector<Batch*> v;
  return v;
}
}  // namespace

bool context_in_batch(int handle) {
  for (const Batch* b : batches())
    if (b)
      for (int h : b->handles)
        if (h == handle) return true;
  return false;
}

namespace {
Batch* batch_of(int id) {
  if (id <= 0 || id > (int)batches().size() || !batches()[id - 1]) {
    fail("no batch %d", id);
    return nullptr;
  }
  return batches()[id - 1];
}
void free_batch(Batch* b) {
  if (b->stream) hipStreamSynchronize(b->stream);
  for (int id : b->graph_ids)
    if (id <= (int)graphs().size() && graphs()[id - 1]) {
      release_graph(graphs()[id - 1]);
      graphs()[id - 1] = nullptr;
    }
  if (b->d_m) hipFree(b->d_m);
  if (b->d_pre) hipFree(b->d_pre);
  if (b->stream) hipStreamDestroy(b->stream);
  delete b;
}
// persistent grid sizes from the device's CU count: 8 blocks of 256 threads, one block of
// seven sweep waves and eight one-wave TRBDF2 blocks resident per CU, times `grid`
void batch_grids(Batch* b) {
  b->t.blocks_thread = 8 * b->cus * b->grid;
  b->t.blocks_sweep = b->cus * b->grid;
  b->t.blocks_diff = 8 * b->cus * b->grid;
}
void batch_launch(Batch* b, hipStream_t st) {
  launch_batch_advection(b->t, b->sweep_pf, st);
  launch_batch_diffusion(b->t, b->diff_chunk, st);
}
}  // namespace

extern "C" {

// Members (contexts with their state resident, residency mode 1) advanced together by
// batched kernels.  Every member must take the fused step's path (all sides global, x3
// not periodic, lx2 and lx3 above 1), use TRBDF2 without the fitting extension, and live
// on one device.  Returns a batch id >= 1.
int gemhip_batch_create(int n, const int* handles) {
  const int keep = g_current_handle;
  if (n <= 0 || !handles) { fail("bad batch arguments"); return -1; }
  auto* b = new Batch();
  b->handles.assign(handles, handles + n);
  for (int m = 0; m < n; ++m) {
    if (gemhip_context_select(handles[m])) { delete b; return -1; }
    Ctx& c = ctx();
    const char* why = nullptr;
    if (!ready(true) || c.residency != 1) why = "is not set up with its state in HBM";
    else if (!preconvert_fusable(c.g) || !postconvert_fusable(c.g))
      why = "needs global sides, no periodic x3 and lx2, lx3 > 1";
    else if (c.diffsolvetype != 2 || c.s.kappa || c.s.qtop)
      why = "must use TRBDF2 without the fitting extension";
    else if (m > 0 && c.device != b->device) why = "is on another device";
    if (why) {                                 // reported in the caller's context
      gemhip_context_select(keep);
      fail("batch member %d %s", m, why);
      delete b;
      return -1;
    }
    if (!ensure_scratch() || !hipok(hipStreamSynchronize(c.stream), "member sync")) {
      const std::string msg = c.last_error;
      gemhip_context_select(keep);
      fail("batch member %d: %s", m, msg.c_str());
      delete b;
      return -1;
    }
    if (m == 0) b->device = c.device;
    c.vsi_current = false;
    b->host.push_back(BatchMember{c.g, c.s, 0.0, c.Teinf, c.d_scratch});
  }
  gemhip_context_select(keep);
  std::vector<long> pre((size_t)BK_COUNT * (n + 1));
  for (int k = 0; k < BK_COUNT; ++k) {
    long acc = 0;
    for (int m = 0; m < n; ++m) {
      pre[(size_t)k * (n + 1) + m] = acc;
      acc += batch_units(b->host[m].g, k);
    }
    pre[(size_t)k * (n + 1) + n] = acc;
    b->t.total[k] = acc;
  }
  bool ok = hipok(hipSetDevice(b->device), "hipSetDevice") &&
            hipok(hipDeviceGetAttribute(&b->cus, hipDeviceAttributeMultiprocessorCount, b->device),
                  "CU count") &&
            hipok(hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking), "batch stream") &&
            hipok(hipMalloc((void**)&b->d_m, n * sizeof(BatchMember)), "batch members") &&
            hipok(hipMalloc((void**)&b->d_pre, pre.size() * sizeof(long)), "batch units") &&
            hipok(hipMemcpy(b->d_m, b->host.data(), n * sizeof(BatchMember), hipMemcpyHostToDevice),
                  "batch members") &&
            hipok(hipMemcpy(b->d_pre, pre.data(), pre.size() * sizeof(long), hipMemcpyHostToDevice),
                  "batch units");
  if (!ok) { free_batch(b); return -1; }
  b->t.m = b->d_m;
  b->t.pre = b->d_pre;
  b->t.n = n;
  batch_grids(b);
  batches().push_back(b);
  return (int)batches().size();
}

// The members' time steps (one per member), read by the batched kernels from HBM: a
// captured batch graph picks up new values without being captured again.
int gemhip_batch_set_dt(int id, const double* dt) {
  Batch* b = batch_of(id);
  if (!b) return 1;
  if (!dt) { fail("null dt"); return 2; }
  for (size_t m = 0; m < b->host.size(); ++m) b->host[m].dt = dt[m];
  if (!hipok(hipStreamSynchronize(b->stream), "batch sync")) return 3;
  return hipok(hipMemcpy(b->d_m, b->host.data(), b->host.size() * sizeof(BatchMember),
                         hipMemcpyHostToDevice), "batch dt") ? 0 : 3;
}

// Kernel variants and the persistent grid: "sweep_pf" (1, 2), "diff_chunk" (4, 8) and
// "grid" (1, 2: resident blocks per CU times this).  Results do not depend on them.
int gemhip_batch_set_tuning(int id, const char* knob, int value) {
  Batch* b = batch_of(id);
  if (!b) return 1;
  const std::string k = knob ? knob : "";
  if (k == "sweep_pf" && (value == 1 || value == 2)) b->sweep_pf = value;
  else if (k == "diff_chunk" && (value == 4 || value == 8)) b->diff_chunk = value;
  else if (k == "grid" && (value == 1 || value == 2)) { b->grid = value; batch_grids(b); }
  else { fail("unknown batch tuning knob or value: %s = %d", k.c_str(), value); return 2; }
  return 0;
}

// One transport step of every member, enqueued on the batch's stream.
int gemhip_batch_step(int id) {
  Batch* b = batch_of(id);
  if (!b) return 1;
  batch_launch(b, b->stream);
  return hipok(hipGetLastError(), "batch launch") ? 0 : 2;
}

int gemhip_batch_synchronize(int id) {
  Batch* b = batch_of(id);
  if (!b) return 1;
  return hipok(hipStreamSynchronize(b->stream), "batch sync") ? 0 : 2;
}

// nsteps batched steps captured into one HIP graph, launched with gemhip_ensemble_launch on
// the batch's stream (so it is ordered with gemhip_batch_step); it holds the members like
// an ensemble graph does and is released with the batch at the latest.  Returns a graph
// id >= 1.
int gemhip_batch_capture(int id, int nsteps) {
  Batch* b = batch_of(id);
  if (!b) return -1;
  if (nsteps <= 0) { fail("bad step count"); return -1; }
  if (!hipok(hipStreamSynchronize(b->stream), "batch sync")) return -1;
  auto* e = new EnsembleGraph();
  e->members = b->handles;
  e->origin = b->stream;
  e->owns_origin = false;
  bool ok = hipok(hipStreamBeginCapture(e->origin, hipStreamCaptureModeGlobal), "begin capture");
  for (int k = 0; ok && k < nsteps; ++k) {
    batch_launch(b, e->origin);
    ok = hipok(hipGetLastError(), "batch launch");
  }
  hipGraph_t graph = nullptr;
  const hipError_t endc = hipStreamEndCapture(e->origin, &graph);
  e->graph = graph;
  ok = ok && hipok(endc, "end capture");
  ok = ok && hipok(hipGraphInstantiate(&e->exec, e->graph, nullptr, nullptr, 0), "instantiate");
  if (!ok) { release_graph(e); return -1; }
  graphs().push_back(e);
  b->graph_ids.push_back((int)graphs().size());
  return (int)graphs().size();
}

int gemhip_batch_release(int id) {
  if (id <= 0 || id > (int)batches().size() || !batches()[id - 1]) return 1;
  free_batch(batches()[id - 1]);
  batches()[id - 1] = nullptr;
  return 0;
}

int gemhip_eval_math(int op, const double* x, double* y, long n) {
  Ctx& c = ctx();
  if (!c.inited) { fail("library not initialised (gemhip_init)"); return 1; }
  if (op < 0 || op > 4 || n < 0 || (n > 0 && (!x || !y))) { fail("bad gemhip_eval_math call"); return 2; }
  if (n == 0) return 0;
  const long nx = op == 4 ? 2 * n : n;               // op 4 reads n pairs
  double *dx = nullptr, *dy = nullptr;
  bool ok = hipok(hipMalloc((void**)&dx, nx * sizeof(double)), "eval alloc") &&
            hipok(hipMalloc((void**)&dy, n * sizeof(double)), "eval alloc") &&
            hipok(hipMemcpy(dx, x, nx * sizeof(double), hipMemcpyDefault), "eval upload");
  if (ok) {
    hipLaunchKernelGGL(k_eval_math, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, c.stream, op,
                       dx, dy, n);
    ok = hipok(hipGetLastError(), "eval launch") &&
         hipok(hipStreamSynchronize(c.stream), "eval sync") &&
         hipok(hipMemcpy(y, dy, n * sizeof(double), hipMemcpyDefault), "eval download");
  }
  if (dx) hipFree(dx);
  if (dy) hipFree(dy);
  return ok ? 0 : 3;
}

}  // extern "C"

// ---- plasma output and initial conditions (io_nompi.f90:183-345, timeutils.f90) ----
namespace gemhip {
namespace {

// timeutils.f90:13-27: days of a month, leap years by the Gregorian rule
int month_days(int year, int month) {
  if (year < 1600 || year > 2500 || month < 1 || month > 12) return -1;
  static const int len[12] = {31, 28, 31, 30, 31, 30, 31, 31, 30, 31, 30, 31};
  const bool leap = (year % 4 == 0 && year % 100 != 0) || year % 400 == 0;
  return month == 2 && leap ? 29 : len[month - 1];
}

// timeutils.f90:110-130: carry days beyond the month into months, months beyond 12 into years
bool carry_date(int& year, int& month, int& day) {
  if (month < 1 || day < 1) return false;
  while (true) {
    year += (month - 1) / 12;
    month = (month - 1) % 12 + 1;
    const int dm = month_days(year, month);
    if (dm < 0) return false;
    if (day <= dm) return true;
    day -= dm;
    ++month;
  }
}

// timeutils.f90:133-196: outdir/YYYYMMDD_SSSSS.ffffff for a double UTsec, rounded to the
// nearest 10 ms (half away from zero, as nint); false where the reference stops.
bool date_stem(const char* outdir, const int* ymd, double UTsec, std::string& out) {
  int year = ymd[0], month = ymd[1], day = ymd[2];
  const double hundredths = std::round(UTsec * 100.0);
  if (!(std::fabs(hundredths) <= 2.0e8)) return false;
  long ms = (long)hundredths * 10;
  long sec = ms / 1000;
  if (sec < 0 || sec > 86400) return false;
  if (sec == 86400) {
    ++day;
    sec = ms = 0;
    if (!carry_date(year, month, day)) return false;
  }
  const long usec = ms * 1000 - sec * 1000000;
  if (year < 0 || year > 9999 || month < 0 || month > 99 || day < 0 || day > 99) return false;
  char buf[64];
  if (usec < 0)   // UTsec in (-1, 0): sec truncates to 0 and the I6.6 field overflows to '*'
    snprintf(buf, sizeof buf, "%04d%02d%02d_%05ld.******", year, month, day, sec);
  else
    snprintf(buf, sizeof buf, "%04d%02d%02d_%05ld.%06ld", year, month, day, sec, usec);
  out = std::string(outdir) + "/" + buf;
  return true;
}

// reader_hdf5.f90:64-92: /lx1 /lx2 /lx3, else /lxs, else /lx (three integers)
bool read_simsize(const char* path, int lin[3]) {
  std::string err;
  H5In f(path, err);
  if (!f.open()) { fail("%s", err.c_str()); return false; }
  long d[8];
  auto scalar = [&](const char* name, int& v) {
    const int r = f.ndims(name, d);
    long cnt = 1;
    for (int k = 0; k < r; ++k) cnt *= d[k];
    int buf[8];
    if (r < 0 || cnt < 1 || cnt > 8 || !f.read(name, true, buf)) return false;
    v = buf[0];
    return true;
  };
  bool got = false;
  if (f.ndims("/lx1", d) >= 0) {
    got = scalar("/lx1", lin[0]) && scalar("/lx2", lin[1]) && scalar("/lx3", lin[2]);
  } else {
    for (const char* name : {"/lxs", "/lx"}) {
      const int r = f.ndims(name, d);
      if (r < 0) continue;
      int buf[8];
      got = r == 1 && d[0] == 3 && f.read(name, true, buf);
      if (got) for (int k = 0; k < 3; ++k) lin[k] = buf[k];
      break;
    }
  }
  if (!got) {
    fail("%s: no grid size (/lx1 /lx2 /lx3, /lxs or /lx)%s%s", path, err.empty() ? "" : ": ",
         err.c_str());
    return false;
  }
  return true;
}

bool ensure_output_buffers(size_t dev_bytes, size_t host_bytes) {
  Ctx& c = ctx();
  if (c.out_bytes < dev_bytes) {
    if (c.d_out) hipFree(c.d_out);
    c.d_out = nullptr;
    c.out_bytes = 0;
    if (!hipok(hipMalloc((void**)&c.d_out, dev_bytes), "output staging")) return false;
    c.out_bytes = dev_bytes;
  }
  if (c.h_out_bytes < host_bytes) {
    if (c.h_out) hipHostFree(c.h_out);
    c.h_out = nullptr;
    c.h_out_bytes = 0;
    if (!hipok(hipHostMalloc((void**)&c.h_out, host_bytes, hipHostMallocDefault),
               "pinned output buffer"))
      return false;
    c.h_out_bytes = host_bytes;
  }
  return true;
}

}  // namespace
}  // namespace gemhip

extern "C" {

int gemhip_register_output(const char* outdir, int flagoutput, int comp_lvl) {
  Ctx& c = ctx();
  if (!outdir) { fail("outdir is a null pointer"); return 1; }
  if (comp_lvl < 0 || comp_lvl > 9) { fail("compression level %d not in 0..9", comp_lvl); return 2; }
  c.outdir = outdir;
  c.flagoutput = flagoutput;
  c.comp_lvl = comp_lvl;
  c.output_ok = true;
  return 0;
}

int gemhip_output_filename(const char* outdir, const int* ymd, double UTsec, int identifier,
                           char* out, long len) {
  if (!outdir || !ymd || !out) { fail("null argument to gemhip_output_filename"); return 1; }
  std::string p;
  if (!date_stem(outdir, ymd, UTsec, p)) {
    fail("timeutils.f90::utsec2filestem did NOT satisfy 0 <= seconds < 86400 (or a bad date)");
    return 2;
  }
  p += "_" + std::to_string(identifier) + ".h5";
  if ((long)p.size() + 1 > len) { fail("file name buffer too short (%ld)", len); return 3; }
  memcpy(out, p.c_str(), p.size() + 1);
  return 0;
}

}  // extern "C"

namespace gemhip {
namespace {
// The output file of the state (both writers of the reference): the patch writer of
// io_nompi.f90:183-268 (the name carries the patch number, the file the limits and
// /patchID; libgemini_c.f90:267-284 always passes them) or, with root, the root writer of
// plasma_output.f90:76-136 and plasma_output_hdf5.f90:10-65 (the plain date name, and TEall
// in place of Teall).  Returns 0 or an error code.
int write_plasma_output(const int* ymd, const double* UTsec, double** fluidvars,
                        double** electrovars, bool root, int patch, const double* const* lims3) {
  if (!ready()) return 1;
  Ctx& c = ctx();
  if (!c.output_ok) { fail("output directory not registered (gemhip_register_output)"); return 2; }
  if (!check_host(electrovars, "electrovars")) return 3;
  if (c.residency == 0 && !check_host(fluidvars, "fluidvars")) return 3;
  std::string path;
  if (!date_stem(c.outdir.c_str(), ymd, *UTsec, path)) {
    fail("timeutils.f90::utsec2filestem did NOT satisfy 0 <= seconds < 86400 (or a bad date)");
    return 4;
  }
  path += root ? std::string(".h5") : "_" + std::to_string(patch) + ".h5";
  std::string err;
  if (!h5_available(err)) { fail("%s", err.c_str()); return 5; }

  const DevGrid& g = c.g;
  const int flag = c.flagoutput;
  const long n = (long)c.lx1 * c.lx2 * c.lx3, m = (long)c.lx2 * c.lx3;
  const long count = output_count(g, flag);
  const size_t ebytes = 4 * (size_t)g.N * sizeof(double);
  if (!ensure_output_buffers(ebytes + count * sizeof(float), count * sizeof(float))) return 6;
  Io io{fluidvars, nullptr, electrovars};
  if (!upload(io, {NS, VS1, VS2, VS3, TS})) return 6;
  // J1 J2 J3 Phi are the contiguous electrovars fields 4-7
  if (!hipok(hipMemcpyAsync(c.d_out, *electrovars + 3 * g.N, ebytes, hipMemcpyDefault, c.stream),
             "electrovars upload"))
    return 6;
  float* packed = reinterpret_cast<float*>(reinterpret_cast<char*>(c.d_out) + ebytes);
  launch_output_pack(g, c.d_fluid, c.d_out, flag, !root, packed, c.stream);
  if (!hipok(hipGetLastError(), "output pack launch")) return 6;

  // field datasets in file order (name, species count, values)
  struct Field { const char* name; int nsp; };
  std::vector<Field> fields;
  if (flag == 2)
    fields = {{"/neall", 0}, {"/v1avgall", 0}, {"/Tavgall", 0}, {root ? "/TEall" : "/Teall", 0}, {"/J1all", 0},
              {"/J2all", 0}, {"/J3all", 0}, {"/v2avgall", 0}, {"/v3avgall", 0}};
  else if (flag == 3)
    fields = {{"/neall", 0}};
  else
    fields = {{"/nsall", 7}, {"/vs1all", 7}, {"/Tsall", 7}, {"/J1all", 0}, {"/J2all", 0},
              {"/J3all", 0}, {"/v2avgall", 0}, {"/v3avgall", 0}};
  const int flagv = flag, id = patch;
  const int ymdv[3] = {ymd[0], ymd[1], ymd[2]};
  const float uthour = (float)(*UTsec / 3600.0);
  float lims[6] = {0, 0, 0, 0, 0, 0};
  for (int k = 0; !root && k < 6; ++k) lims[k] = (float)lims3[k / 2][k % 2];
  std::vector<H5Out> ds = {
      {"/flagoutput", H5Out::I32, 0, {0}, &flagv, nullptr},
      {"/time/ymd", H5Out::I32, 1, {3}, ymdv, nullptr},
      {"/time/UThour", H5Out::F32, 0, {0}, &uthour, nullptr}};
  if (!root) {
    ds.push_back({"/x1lims", H5Out::F32, 1, {2}, lims, nullptr});
    ds.push_back({"/x2lims", H5Out::F32, 1, {2}, lims + 2, nullptr});
    ds.push_back({"/x3lims", H5Out::F32, 1, {2}, lims + 4, nullptr});
    ds.push_back({"/patchID", H5Out::I32, 0, {0}, &id, nullptr});
  }
  // one device-to-host copy per dataset, each followed by an event, so the file write of
  // one dataset overlaps the copy of the next
  std::vector<hipEvent_t> evs;
  long off = 0;
  bool ok = true;
  auto add = [&](const char* name, int rank, std::initializer_list<long> dims, long len) {
    hipEvent_t e = take_event();
    evs.push_back(e);
    ok = ok && hipok(hipMemcpyAsync(c.h_out + off, packed + off, len * sizeof(float),
                                    hipMemcpyDeviceToHost, c.stream), "output download") &&
         hipok(hipEventRecord(e, c.stream), "output event");
    H5Out d{name, H5Out::F32, rank, {0, 0, 0, 0}, c.h_out + off, e};
    int k = 0;
    for (long v : dims) d.dims[k++] = v;
    ds.push_back(d);
    off += len;
  };
  for (const Field& f : fields) {
    if (f.nsp) add(f.name, 4, {f.nsp, c.lx3, c.lx2, c.lx1}, f.nsp * n);
    else add(f.name, 3, {c.lx3, c.lx2, c.lx1}, n);
  }
  add("/Phiall", 2, {c.lx3, c.lx2}, m);
  if (ok && !h5_write(path.c_str(), ds.data(), (int)ds.size(), c.comp_lvl, err)) {
    fail("plasma output %s: %s", path.c_str(), err.c_str());
    ok = false;
  }
  hipStreamSynchronize(c.stream);
  for (hipEvent_t e : evs) c.pool.push_back(e);
  return ok ? 0 : 7;
}
}  // namespace
}  // namespace gemhip

extern "C" {

// timeutils.f90:79-107 (dateinc, bound by libgemini_c.f90:1144-1150): advance the date and
// UT seconds by dt, carrying whole days into the date
void dateinc_C(double* dt, int* ymd, double* UTsec) {
  if (!dt || !ymd || !UTsec) { fail("null argument to dateinc_C"); return; }
  if (ymd[2] < 1) { fail("temporal:timeutils:dateinc(): days are positive integers"); return; }
  if (*UTsec < 0) { fail("negative UTsec, simulation should go forward in time only!"); return; }
  if (*dt < 0) { fail("negative dtsec, simulation should go forward in time only!"); return; }
  if (*dt > 86400) {
    fail("excessively large dtsec > 86400, simulation step should be small enough!");
    return;
  }
  int year = ymd[0], month = ymd[1], day = ymd[2];
  double ut = *UTsec + *dt;
  while (ut >= 86400) {
    ut = ut - 86400.0;
    ++day;
    if (!carry_date(year, month, day)) { fail("day_wrap: months and days must be positive"); return; }
  }
  ymd[0] = year;
  ymd[1] = month;
  ymd[2] = day;
  *UTsec = ut;
}

// the name of the root writer's file: outdir/YYYYMMDD_SSSSS.ffffff.h5 (plasma_output_hdf5.f90:17)
int gemhip_output_filename_root(const char* outdir, const int* ymd, double UTsec, char* out,
                                long len) {
  if (!outdir || !ymd || !out) { fail("null argument to gemhip_output_filename_root"); return 1; }
  std::string p;
  if (!date_stem(outdir, ymd, UTsec, p)) {
    fail("timeutils.f90::utsec2filestem did NOT satisfy 0 <= seconds < 86400 (or a bad date)");
    return 2;
  }
  p += ".h5";
  if ((long)p.size() + 1 > len) { fail("file name buffer too short (%ld)", len); return 3; }
  memcpy(out, p.c_str(), p.size() + 1);
  return 0;
}

// io_nompi.f90:183-268 through libgemini_c.f90:267-284
void plasma_output_nompi_C(void** cfgC, int* ymd, double* UTsec, double** fluidvars,
                           double** electrovars, int* identifier, double* x1lims,
                           double* x2lims, double* x3lims) {
  (void)cfgC;
  if (!ymd || !UTsec || !identifier || !x1lims || !x2lims || !x3lims) {
    fail("null argument to plasma_output_nompi_C");
    return;
  }
  const double* lims3[3] = {x1lims, x2lims, x3lims};
  write_plasma_output(ymd, UTsec, fluidvars, electrovars, false, *identifier, lims3);
}

// output_plasma (plasma_output.f90:23-31, 76-136 -> plasma_output_hdf5.f90:10-65): the file
// gemini.bin writes every output time; one process holds the whole grid, so the root's
// gather is the identity
int gemhip_output_plasma(const int* ymd, double UTsec, double** fluidvars, double** electrovars) {
  if (!ymd) { fail("null argument to gemhip_output_plasma"); return 1; }
  return write_plasma_output(ymd, &UTsec, fluidvars, electrovars, true, 0, nullptr);
}


// reader_hdf5.f90:64-92 (get_simsize3): the full-grid sizes of a simsize file
int gemhip_read_simsize(const char* path, int* lx1, int* lx2, int* lx3) {
  if (!path || !lx1 || !lx2 || !lx3) { fail("null argument to gemhip_read_simsize"); return 1; }
  int lin[3];
  if (!read_simsize(path, lin)) return 2;
  *lx1 = lin[0];
  *lx2 = lin[1];
  *lx3 = lin[2];
  return 0;
}

// readgrid_hdf5.f90:11-39 (get_grid3_coords_hdf5): the ghosted coordinates x1(-1:lx1+2),
// x2(-1:lx2+2), x3(-1:lx3+2) and the grid centre of path, or of path/simgrid.h5 when path
// does not name a simgrid.h5 file
int gemhip_read_grid_coords(const char* path, int lx1, int lx2, int lx3, double* x1, double* x2,
                            double* x3, double* glonctr, double* glatctr) {
  if (!path || !x1 || !x2 || !x3 || !glonctr || !glatctr) {
    fail("null argument to gemhip_read_grid_coords");
    return 1;
  }
  std::string fn = path;
  if (fn.find("simgrid.h5") == std::string::npos) fn += "/simgrid.h5";
  std::string err;
  H5In f(fn.c_str(), err);
  if (!f.open()) { fail("%s", err.c_str()); return 2; }
  const struct { const char* name; double* dst; int n; } xs[3] = {
      {"/x1", x1, lx1 + 4}, {"/x2", x2, lx2 + 4}, {"/x3", x3, lx3 + 4}};
  for (const auto& x : xs) {
    long d[8];
    const int r = f.ndims(x.name, d);
    if (r != 1 || d[0] != x.n) {
      fail("%s: %s must hold %d values (the cells and two ghost cells at each end)", fn.c_str(),
           x.name, x.n);
      return 3;
    }
    if (!f.read(x.name, false, x.dst)) { fail("%s: %s", fn.c_str(), err.c_str()); return 4; }
  }
  for (const char* name : {"/glonctr", "/glatctr"}) {
    long d[8];
    const int r = f.ndims(name, d);
    if (r < 0) {
      fail("please specify /glonctr and /glatctr in %s", fn.c_str());
      return 5;
    }
    long cnt = 1;
    for (int k = 0; k < r; ++k) cnt *= d[k];
    if (cnt != 1) { fail("%s: %s must be one value", fn.c_str(), name); return 5; }
  }
  if (!f.read("/glonctr", false, glonctr) || !f.read("/glatctr", false, glatctr)) {
    fail("%s: %s", fn.c_str(), err.c_str());
    return 6;
  }
  return 0;
}

// io_nompi.f90:271-345 with the zeroing of plasma_input_hdf5.f90:78-86: sizes from
// indatsize (/lx1 /lx2 /lx3, else /lxs, else /lx) must equal lx1, lx2, lx3; ns, vs1 and Ts
// of fluidvars are zeroed and their interiors read from /nsall /vs1all /Tsall of indatfile;
// the potential field of electrovars (if given) is zeroed and each interior x1 plane set to
// /Phiall (0 if absent).  Host blocks only: with the HBM copy resident, push afterwards.
int gemhip_read_initial_conditions(const char* indatsize, const char* indatfile, int lx1,
                                   int lx2, int lx3, double* fluidvars, double* electrovars) {
  if (!indatsize || !indatfile || !fluidvars) { fail("null argument to gemhip_read_initial_conditions"); return 1; }
  if (lx1 < 1 || lx2 < 1 || lx3 < 1) { fail("grid size %d x %d x %d", lx1, lx2, lx3); return 1; }
  std::string err;
  int lin[3] = {0, 0, 0};
  if (!read_simsize(indatsize, lin)) return 3;
  if (lin[0] != lx1 || lin[1] != lx2 || lin[2] != lx3) {
    fail("The input data must be the same size as the grid which you are running the simulation "
         "on - use a script to interpolate up/down to the simulation grid (file %d x %d x %d, "
         "grid %d x %d x %d)", lin[0], lin[1], lin[2], lx1, lx2, lx3);
    return 4;
  }
  const long L1 = lx1 + 4, L2 = lx2 + 4, L3 = lx3 + 4, N = L1 * L2 * L3, B = 7 * N;
  H5In f(indatfile, err);
  if (!f.open()) { fail("%s", err.c_str()); return 5; }
  const long mdims[4] = {7, L3, L2, L1}, start[4] = {0, 2, 2, 2};
  const struct { const char* name; double* dst; } vars[3] = {
      {"/nsall", fluidvars}, {"/vs1all", fluidvars + B}, {"/Tsall", fluidvars + 4 * B}};
  for (const auto& v : vars) {
    long d[8];
    const int r = f.ndims(v.name, d);
    if (r != 4 || d[0] != 7 || d[1] != lx3 || d[2] != lx2 || d[3] != lx1) {
      fail("%s: %s must have the shape (lx1, lx2, lx3, 7) = (%d, %d, %d, 7)", indatfile, v.name,
           lx1, lx2, lx3);
      return 6;
    }
    memset(v.dst, 0, B * sizeof(double));
    if (!f.read(v.name, false, v.dst, 4, mdims, start)) { fail("%s: %s", indatfile, err.c_str()); return 7; }
  }
  if (!electrovars) return 0;
  std::vector<double> slab((size_t)lx2 * lx3, 0.0);
  long d[8];
  const int r = f.ndims("/Phiall", d);
  if (r == 1) {
    // a 1-D potential spans x3 when lx2 = 1 and x2 when lx3 = 1 (io_nompi.f90:318-333)
    const long want = lx2 == 1 ? lx3 : lx2;
    if (d[0] != want || (lx2 != 1 && lx3 != 1)) {
      fail("%s: a 1-D /Phiall needs lx2 = 1 or lx3 = 1 and their other size (%ld values)",
           indatfile, d[0]);
      return 8;
    }
  } else if (r == 2) {
    if (d[0] != lx3 || d[1] != lx2) {
      fail("%s: /Phiall must have the shape (lx2, lx3) = (%d, %d)", indatfile, lx2, lx3);
      return 8;
    }
  } else if (r >= 0) {
    fail("%s: /Phiall has %d dimensions", indatfile, r);
    return 8;
  }
  if (r >= 0 && !f.read("/Phiall", false, slab.data())) { fail("%s: %s", indatfile, err.c_str()); return 9; }
  double* Phi = electrovars + 6 * N;
  memset(Phi, 0, N * sizeof(double));
  for (long i3 = 0; i3 < lx3; ++i3)
    for (long i2 = 0; i2 < lx2; ++i2) {
      const double v = slab[i2 + lx2 * i3];
      double* row = Phi + 2 + L1 * ((i2 + 2) + L2 * (i3 + 2));
      for (long i1 = 0; i1 < lx1; ++i1) row[i1] = v;
    }
  return 0;
}

}  // extern "C"
