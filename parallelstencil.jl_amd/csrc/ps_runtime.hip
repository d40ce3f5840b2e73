// libpsamd runtime: the C-ABI declared in include/ps_amd.h.
//
// Device memory, streams and events are plain HIP; generated kernels are
// compiled with hiprtc for gfx950 and cached on disk by the hash of their
// source and compile options, so a kernel compiles once per machine image.
#include <hip/hip_runtime.h>
#include <hip/hiprtc.h>

#include <algorithm>
#include <atomic>
#include <cerrno>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <dlfcn.h>
#include <fstream>
#include <map>
#include <memory>
#include <unordered_map>
#include <mutex>
#include <sstream>
#include <string>
#include <sys/stat.h>
#include <unistd.h>
#include <vector>

#include "ps_amd.h"
#include "ps_lower.h"

using namespace ps;

// ---------------------------------------------------------------------------
// error state
// ---------------------------------------------------------------------------
static thread_local std::string g_err;

static int fail(int code, const std::string& msg) {
  g_err = msg;
  return code;
}

#define PS_HIP(call)                                                                       \
  do {                                                                                     \
    hipError_t e_ = (call);                                                                \
    if (e_ != hipSuccess) return fail(PS_EHIP, std::string(#call) + ": " + hipGetErrorString(e_)); \
  } while (0)

#define PS_GUARD_BEGIN try {
#define PS_GUARD_END                                       \
  }                                                        \
  catch (const PsError& e) {                               \
    return fail(e.code, e.what());                         \
  }                                                        \
  catch (const std::exception& e) {                        \
    return fail(PS_EARGUMENT, e.what());                   \
  }

extern "C" const char* ps_last_error(void) { return g_err.c_str(); }

// extra row padding (bytes) applied to fields whose row stride is a multiple of 4 KiB
static int64_t g_pitch_pad = [] {
  const char* e = std::getenv("PS_AMD_PITCH_PAD");
  return e ? (int64_t)std::atoll(e) : (int64_t)0;
}();
extern "C" int ps_version(void) { return 10000; }

// ---------------------------------------------------------------------------
// compiled-in helper kernels (fill, copy, reductions, error flag)
// ---------------------------------------------------------------------------
template <class T>
__global__ void __launch_bounds__(256) ps_fill_k(T* p, long long nx, long long ny, long long nz, long long sx, long long sxy, T v) {
  long long i = (long long)blockIdx.x * 256 + threadIdx.x;
  long long n = nx * ny * nz;
  for (; i < n; i += (long long)gridDim.x * 256) {
    long long x = i % nx, y = (i / nx) % ny, z = i / (nx * ny);
    p[x + y * sx + z * sxy] = v;
  }
}

template <class T>
__global__ void __launch_bounds__(256) ps_copy_k(T* __restrict__ d, const T* __restrict__ s, long long nx, long long ny, long long nz,
                                                 long long dsx, long long dsxy, long long ssx, long long ssxy) {
  long long i = (long long)blockIdx.x * 256 + threadIdx.x;
  long long n = nx * ny * nz;
  for (; i < n; i += (long long)gridDim.x * 256) {
    long long x = i % nx, y = (i / nx) % ny, z = i / (nx * ny);
    d[x + y * dsx + z * dsxy] = s[x + y * ssx + z * ssxy];
  }
}

__global__ void ps_set_i32_k(int* p, int v) { *p = v; }

// max that keeps a NaN once one has been seen, as Julia's maximum(abs.(R)) does (fmax drops it)
__device__ __forceinline__ double ps_nanmax(double a, double b) { return (a != a || b != b) ? a + b : fmax(a, b); }

template <class T>
__device__ __forceinline__ double ps_to_f64(T v) { return (double)v; }
template <>
__device__ __forceinline__ double ps_to_f64<unsigned short>(unsigned short v) { return (double)__uint_as_float(((unsigned)v) << 16); }

// Stage 1: per-block partial (sum, sum|x|, max|x|, sum x^2) over the logical box in a
// fixed order, so repeated reductions of the same data give the same bits.
template <class T>
__global__ void __launch_bounds__(256) ps_reduce1_k(const T* __restrict__ p, long long nx, long long ny, long long nz, long long sx,
                                                    long long sxy, double* __restrict__ part) {
  __shared__ double s0[256], s1[256], s2[256], s3[256];
  long long n = nx * ny * nz;
  double a = 0.0, b = 0.0, c = 0.0, d = 0.0;
  for (long long i = (long long)blockIdx.x * 256 + threadIdx.x; i < n; i += (long long)gridDim.x * 256) {
    long long x = i % nx, y = (i / nx) % ny, z = i / (nx * ny);
    double v = ps_to_f64<T>(p[x + y * sx + z * sxy]);
    a += v;
    b += fabs(v);
    c = ps_nanmax(c, fabs(v));
    d += v * v;
  }
  s0[threadIdx.x] = a;
  s1[threadIdx.x] = b;
  s2[threadIdx.x] = c;
  s3[threadIdx.x] = d;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) {
      s0[threadIdx.x] += s0[threadIdx.x + w];
      s1[threadIdx.x] += s1[threadIdx.x + w];
      s2[threadIdx.x] = ps_nanmax(s2[threadIdx.x], s2[threadIdx.x + w]);
      s3[threadIdx.x] += s3[threadIdx.x + w];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    part[4 * blockIdx.x + 0] = s0[0];
    part[4 * blockIdx.x + 1] = s1[0];
    part[4 * blockIdx.x + 2] = s2[0];
    part[4 * blockIdx.x + 3] = s3[0];
  }
}

__global__ void __launch_bounds__(256) ps_reduce2_k(const double* __restrict__ part, int nparts, double* __restrict__ out) {
  __shared__ double s0[256], s1[256], s2[256], s3[256];
  double a = 0.0, b = 0.0, c = 0.0, d = 0.0;
  for (int i = threadIdx.x; i < nparts; i += 256) {
    a += part[4 * i];
    b += part[4 * i + 1];
    c = ps_nanmax(c, part[4 * i + 2]);
    d += part[4 * i + 3];
  }
  s0[threadIdx.x] = a;
  s1[threadIdx.x] = b;
  s2[threadIdx.x] = c;
  s3[threadIdx.x] = d;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) {
      s0[threadIdx.x] += s0[threadIdx.x + w];
      s1[threadIdx.x] += s1[threadIdx.x + w];
      s2[threadIdx.x] = ps_nanmax(s2[threadIdx.x], s2[threadIdx.x + w]);
      s3[threadIdx.x] += s3[threadIdx.x + w];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    out[0] = s0[0];
    out[1] = s1[0];
    out[2] = s2[0];
    out[3] = s3[0];
  }
}

// Fields reduced by one call of ps_field_reduce_noalloc, and its scratch: the stage-1 partials of
// each field (kReduceParts blocks x 4 statistics), 4 x 2048 x 4 x 8 B = 256 KiB per GPU. Each GPU
// the library runs on holds its own copy until the process ends.
constexpr int kNoallocFields = 4;
constexpr int kReduceParts = 2048;
static __device__ __align__(256) double scratch[kNoallocFields * kReduceParts * 4];

struct ps_red_meta {
  int nparts[kNoallocFields];
  long long n[kNoallocFields];
};

// Block b combines the partials of field b exactly as ps_reduce2_k does, then writes the requested
// statistic (op as in ps_field_reduce) to out[b].
__global__ void __launch_bounds__(256) ps_reduce2_op_k(const double* __restrict__ part, ps_red_meta meta, int op,
                                                       double* __restrict__ out) {
  __shared__ double s0[256], s1[256], s2[256], s3[256];
  const double* pb = part + (size_t)blockIdx.x * kReduceParts * 4;
  const int nparts = meta.nparts[blockIdx.x];
  double a = 0.0, b = 0.0, c = 0.0, d = 0.0;
  for (int i = threadIdx.x; i < nparts; i += 256) {
    a += pb[4 * i];
    b += pb[4 * i + 1];
    c = ps_nanmax(c, pb[4 * i + 2]);
    d += pb[4 * i + 3];
  }
  s0[threadIdx.x] = a;
  s1[threadIdx.x] = b;
  s2[threadIdx.x] = c;
  s3[threadIdx.x] = d;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) {
      s0[threadIdx.x] += s0[threadIdx.x + w];
      s1[threadIdx.x] += s1[threadIdx.x + w];
      s2[threadIdx.x] = ps_nanmax(s2[threadIdx.x], s2[threadIdx.x + w]);
      s3[threadIdx.x] += s3[threadIdx.x + w];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    const long long n = meta.n[blockIdx.x];
    double r;
    if (op == 0) r = s0[0];
    else if (op == 1) r = s1[0];
    else if (op == 2) r = s2[0];
    else if (op == 3) r = n ? s1[0] / (double)n : 0.0;
    else r = sqrt(s3[0]);
    out[blockIdx.x] = r;
  }
}

// ---------------------------------------------------------------------------
// global state
// ---------------------------------------------------------------------------
namespace {

struct Module {
  hipModule_t mod = nullptr;
  hipFunction_t fn = nullptr;
  Emitted em;
};

struct LaunchEntry;

struct KernelObj {
  Kernel K;
  std::mutex mu;
  std::map<std::string, std::shared_ptr<Module>> variants;  // key: kinds+variant
  // launch plans (geometry, variant, module) by argument signature: repeated launches of the
  // same shapes skip the planning, the bounds proof and the variant lookup
  std::mutex lmu;
  std::unordered_map<std::string, std::shared_ptr<LaunchEntry>> launches;
  // padding=true kernels comparing parent sizes: the kernel per pattern of comparison results
  std::mutex pmu;
  std::map<std::string, std::shared_ptr<Kernel>> padspec;
};

struct State {
  std::mutex mu;
  bool inited = false;
  int device = -1;
  std::vector<hipStream_t> streams;       // pool 0
  std::vector<hipStream_t> prio_streams;  // pool 1
  hipStream_t stream0 = nullptr;
  int* d_err = nullptr;
  double* d_part = nullptr;
  double* d_out = nullptr;
  double* h_out = nullptr;
  int nparts_cap = 0;
  std::string cache_dir;
  // device address of `scratch` per device (hipGetSymbolAddress at ps_init, that device current)
  std::map<int, double*> scratch_ptr;
};

State& S() {
  static State s;
  return s;
}

size_t elsize(int dtype) {
  switch (dtype) {
    case PS_F64: return 8;
    case PS_F32: return 4;
    case PS_BF16: return 2;
    case PS_I32: return 4;
    case PS_I64: return 8;
    case PS_BOOL: return 1;
  }
  return 0;
}

std::string default_cache_dir() {
  if (const char* e = std::getenv("PS_AMD_CACHE")) return e;
  Dl_info info;
  if (dladdr((void*)&ps_last_error, &info) && info.dli_fname) {
    std::string p = info.dli_fname;
    size_t s = p.rfind('/');
    if (s != std::string::npos) return p.substr(0, s) + "/kcache";
  }
  return "kcache";
}

const std::string& cache_dir() {
  State& s = S();
  if (s.cache_dir.empty()) s.cache_dir = default_cache_dir();
  return s.cache_dir;
}

unsigned long long fnv1a(const std::string& x, unsigned long long h = 1469598103934665603ULL) {
  for (unsigned char c : x) {
    h ^= c;
    h *= 1099511628211ULL;
  }
  return h;
}

std::string hash_hex(const std::string& x) {
  char b[40];
  std::snprintf(b, sizeof b, "%016llx%016llx", fnv1a(x), fnv1a(x, 0x84222325cbf29ce4ULL));
  return b;
}

void mkdirs(const std::string& d) {
  std::string cur;
  std::stringstream ss(d);
  std::string part;
  if (!d.empty() && d[0] == '/') cur = "/";
  while (std::getline(ss, part, '/')) {
    if (part.empty()) continue;
    cur += part + "/";
    mkdir(cur.c_str(), 0755);
  }
}

std::vector<std::string> compile_opts(const Kernel& K) {
  std::vector<std::string> o = {"--offload-arch=gfx950", "-O3", "-std=c++17", "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"};
  o.push_back((K.opt.fp_contract || K.opt.fastmath) ? "-ffp-contract=fast" : "-ffp-contract=off");
  return o;
}

// Compile (or fetch from the disk cache) the code object of an emitted kernel.
std::vector<char> code_object(const Kernel& K, const Emitted& em) {
  auto opts = compile_opts(K);
  std::string key = em.source;
  for (auto& x : opts) key += "\n//" + x;
  int vmaj = 0, vmin = 0;
  hiprtcVersion(&vmaj, &vmin);
  key += "\n//hiprtc " + std::to_string(vmaj) + "." + std::to_string(vmin);
  std::string path = cache_dir() + "/" + hash_hex(key) + ".hsaco";
  {
    std::ifstream f(path, std::ios::binary);
    if (f) {
      std::vector<char> buf((std::istreambuf_iterator<char>(f)), std::istreambuf_iterator<char>());
      if (!buf.empty()) return buf;
    }
  }
  hiprtcProgram prog;
  if (hiprtcCreateProgram(&prog, em.source.c_str(), (K.def.name + ".hip").c_str(), 0, nullptr, nullptr) != HIPRTC_SUCCESS)
    throw PsError(PS_ECOMPILE, "hiprtcCreateProgram failed");
  std::vector<const char*> copts;
  for (auto& x : opts) copts.push_back(x.c_str());
  hiprtcResult r = hiprtcCompileProgram(prog, (int)copts.size(), copts.data());
  if (r != HIPRTC_SUCCESS) {
    size_t ls = 0;
    hiprtcGetProgramLogSize(prog, &ls);
    std::string log(ls + 1, '\0');
    hiprtcGetProgramLog(prog, &log[0]);
    hiprtcDestroyProgram(&prog);
    throw PsError(PS_ECOMPILE, "hiprtc failed to compile kernel " + K.def.name + ":\n" + log + "\n--- source ---\n" + em.source);
  }
  size_t cs = 0;
  hiprtcGetCodeSize(prog, &cs);
  std::vector<char> code(cs);
  hiprtcGetCode(prog, code.data());
  hiprtcDestroyProgram(&prog);
  mkdirs(cache_dir());
  std::string tmp = path + ".tmp" + std::to_string(getpid());
  {
    std::ofstream f(tmp, std::ios::binary);
    f.write(code.data(), (std::streamsize)code.size());
  }
  std::rename(tmp.c_str(), path.c_str());
  return code;
}

// The data array of a CellArray as a dense 1-D field (h2d, d2h, fill, copy and reductions move
// all values of all cells); a plain field is returned unchanged.
ps_field data_of(const ps_field& f) {
  if (f.ncomp <= 0) return f;
  ps_field d = f;
  const int64_t N = f.dims[0] * f.dims[1] * f.dims[2];
  const int64_t B = std::max<int64_t>(f.blocklen, 1);
  const int64_t total = (N + B - 1) / B * B * f.ncomp;
  d.dims[0] = total;
  d.dims[1] = d.dims[2] = 1;
  d.pitch = d.plane = std::max<int64_t>(total, 1);
  d.ndims = 1;
  d.ncomp = 0;
  return d;
}

// The parent of a view (ps_field_view) and the field itself otherwise: extent per dimension,
// offset of the field's first element in it, and the address of the parent's first element.
int64_t pdim(const ps_field& f, int d) { return f.pdims[0] ? f.pdims[d] : f.dims[d]; }
int64_t voffset(const ps_field& f, int d) { return f.pdims[0] ? f.voff[d] : 0; }
void* parent_ptr(const ps_field& f) {
  if (!f.pdims[0] || !f.ptr) return f.ptr;
  return (char*)f.ptr - (f.voff[0] + f.voff[1] * f.pitch + f.voff[2] * f.plane) * (int64_t)elsize(f.dtype);
}
ps_field parent_of(const ps_field& f) {
  ps_field p = f;
  p.ptr = parent_ptr(f);
  for (int d = 0; d < 3; ++d) {
    p.dims[d] = pdim(f, d);
    p.voff[d] = 0;
    p.pdims[d] = 0;
  }
  return p;
}

// Bytes a field argument spans from its first element (aliasing analysis).
int64_t extent_bytes(const ps_field& f) {
  const ps_field d = data_of(f);
  return d.plane * std::max<int64_t>(d.dims[2], 1) * (int64_t)elsize(d.dtype);
}

// 1: one block holds every cell (blocklen >= cells), 2: cells stored whole (blocklen 1),
// 3: blocks of blocklen cells
int cell_layout(const ps_field& f) {
  const int64_t N = f.dims[0] * f.dims[1] * f.dims[2];
  if (f.blocklen >= N) return 1;
  if (f.blocklen == 1) return 2;
  return 3;
}

std::vector<ArgKind> arg_kinds(const ps_arg* args, int nargs) {
  std::vector<ArgKind> k(nargs);
  for (int i = 0; i < nargs; ++i) {
    if (args[i].kind == PS_ARG_FIELD) {
      k[i].is_field = true;
      k[i].dtype = args[i].field.dtype;
      k[i].ndims = args[i].field.ndims;
      if (args[i].field.ncomp > 0) {
        const ps_field& f = args[i].field;
        if (f.pitch != f.dims[0] || f.plane != f.dims[0] * f.dims[1] || f.blocklen < 1)
          throw PsError(PS_EARGUMENT, "argument " + std::to_string(i + 1) + ": a CellArray must be dense with blocklen >= 1");
        k[i].ncomp = f.ncomp;
        for (int d = 0; d < 4; ++d) k[i].celldims[d] = f.celldims[d];
        k[i].cell_layout = cell_layout(f);
        k[i].cellnames = f.cellnames ? f.cellnames : "";
      }
    } else if (args[i].kind == PS_ARG_SCALAR) {
      k[i].dtype = args[i].dtype;
    } else {
      throw PsError(PS_EARGUMENT, "argument " + std::to_string(i + 1) + " has an unknown kind");
    }
    if (elsize(k[i].dtype) == 0) throw PsError(PS_EARGUMENT, "argument " + std::to_string(i + 1) + " has an unsupported type");
  }
  return k;
}

std::string kinds_key(const std::vector<ArgKind>& k) {
  std::string s;
  for (auto& a : k) {
    s += (a.is_field ? "F" : "S") + std::to_string(a.dtype) + (a.is_field ? std::to_string(a.ndims) : "");
    if (a.ncomp > 0)
      s += "c" + std::to_string(a.celldims[0]) + "x" + std::to_string(a.celldims[1]) + "x" + std::to_string(a.celldims[2]) + "x" +
           std::to_string(a.celldims[3]) + "l" +
           std::to_string(a.cell_layout) + "[" + a.cellnames + "]";
    s += ",";
  }
  return s;
}

// The kernel a launch with these arguments runs: the definition itself, or for padding=true
// kernels whose indices compare parent sizes, its specialisation to the arguments' parents
// (*bits receives the comparison results, "" when there are none).
const Kernel& kernel_for(KernelObj* ko, const ps_arg* args, int nargs, std::string* bits) {
  const Kernel& K = ko->K;
  bits->clear();
  if (K.pgt_terms.empty()) return K;
  for (auto& t : K.pgt_terms) {
    if (t[0] >= nargs || t[1] >= nargs || args[t[0]].kind != PS_ARG_FIELD || args[t[1]].kind != PS_ARG_FIELD)
      throw PsError(PS_EARGUMENT, "kernel " + K.def.name + ": padding=true needs array arguments where the kernel reads staggered arrays");
    bits->push_back(pdim(args[t[0]].field, t[2] - 1) > pdim(args[t[1]].field, t[2] - 1) ? '1' : '0');
  }
  std::lock_guard<std::mutex> lk(ko->pmu);
  auto& slot = ko->padspec[*bits];
  if (!slot) slot = std::make_shared<Kernel>(specialize_padding(K, *bits));
  return *slot;
}

std::shared_ptr<Module> get_module(KernelObj* ko, const Kernel& K, const std::vector<ArgKind>& kinds, const Variant& v, bool load) {
  std::string key = kinds_key(kinds) + v.key();
  std::lock_guard<std::mutex> lk(ko->mu);
  auto it = ko->variants.find(key);
  if (it != ko->variants.end() && (!load || it->second->fn)) return it->second;
  auto m = std::make_shared<Module>();
  m->em = lower_kernel(K, kinds, v);
  std::vector<char> code = code_object(K, m->em);
  if (load) {
    hipError_t e = hipModuleLoadData(&m->mod, code.data());
    if (e != hipSuccess) throw PsError(PS_EHIP, std::string("hipModuleLoadData: ") + hipGetErrorString(e));
    e = hipModuleGetFunction(&m->fn, m->mod, m->em.entry.c_str());
    if (e != hipSuccess) throw PsError(PS_EHIP, std::string("hipModuleGetFunction: ") + hipGetErrorString(e));
  }
  ko->variants[key] = m;
  return m;
}

int require_init() {
  if (!S().inited) return fail(PS_ENOTINIT, "ps_init has not been called (no device initialised)");
  return 0;
}

hipStream_t stream_of(void* s) { return s ? (hipStream_t)s : S().stream0; }

}  // namespace

// helpers shared with the other translation units of the library (csrc/ps_internal.h)
int ps_internal_fail(int code, const std::string& msg) { return fail(code, msg); }
int ps_internal_require_init() { return require_init(); }
hipStream_t ps_internal_stream(void* s) { return stream_of(s); }

// ---------------------------------------------------------------------------
// device
// ---------------------------------------------------------------------------
extern "C" int ps_init(int device) {
  State& s = S();
  std::lock_guard<std::mutex> lk(s.mu);
  if (s.inited) {
    if (device != s.device) return fail(PS_EINCOHERENT, "ps_init: already initialised on another device");
    return 0;
  }
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess || n <= 0) return fail(PS_ENODEVICE, "no HIP device available");
  if (device < 0 || device >= n) return fail(PS_EARGUMENT, "ps_init: invalid device index");
  PS_HIP(hipSetDevice(device));
  hipDeviceProp_t p;
  PS_HIP(hipGetDeviceProperties(&p, device));
  if (std::strncmp(p.gcnArchName, "gfx950", 6) != 0)
    return fail(PS_ENODEVICE, std::string("this backend targets gfx950 (MI355X); found ") + p.gcnArchName);
  PS_HIP(hipStreamCreateWithFlags(&s.stream0, hipStreamNonBlocking));
  PS_HIP(hipMalloc(&s.d_err, 256));
  PS_HIP(hipMemset(s.d_err, 0, 256));
  PS_HIP(hipHostMalloc(&s.h_out, 4 * sizeof(double), hipHostMallocDefault));
  PS_HIP(hipMalloc(&s.d_out, 4 * sizeof(double)));
  PS_HIP(hipMalloc(&s.d_part, sizeof(double) * 4 * kReduceParts));
  s.nparts_cap = kReduceParts;
  {
    // the array's name is a host-side stand-in: kernels get this device address instead
    double* sp = nullptr;
    hipError_t e = hipGetSymbolAddress((void**)&sp, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !sp) return fail(PS_EHIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
    s.scratch_ptr[device] = sp;
  }
  s.device = device;
  s.inited = true;
  return 0;
}

extern "C" int ps_device_info(ps_devinfo* out) {
  if (int r = require_init()) return r;
  hipDeviceProp_t p;
  PS_HIP(hipGetDeviceProperties(&p, S().device));
  std::memset(out, 0, sizeof *out);
  std::snprintf(out->name, sizeof out->name, "%s", p.name);
  std::snprintf(out->arch, sizeof out->arch, "%s", p.gcnArchName);
  out->cu_count = p.multiProcessorCount;
  out->wave_size = p.warpSize;
  out->lds_bytes_per_block = (int64_t)p.sharedMemPerBlock;
  out->total_mem_bytes = (int64_t)p.totalGlobalMem;
  out->l2_bytes = p.l2CacheSize;
  out->clock_khz = p.clockRate;
  out->mem_clock_khz = p.memoryClockRate;
  out->mem_bus_width = p.memoryBusWidth;
  // HBM3/HBM3E parts report a memory clock of a quarter of the per-pin data rate: MI300X reports
  // 1.3 GHz on 8192 bits (5.3 TB/s published), MI355X 2.0 GHz on 8192 bits (8 TB/s published)
  out->hbm_peak_gbs = 4.0 * (double)p.memoryClockRate * 1e3 * (double)p.memoryBusWidth / 8.0 / 1e9;
  return 0;
}

// Set once a bounds-checked variant has been launched (only those write the error flag): until
// then a synchronisation needs no device-to-host read of the flag.
static std::atomic<bool> g_err_armed{false};

static int check_err_flag(hipStream_t st) {
  if (!g_err_armed.load(std::memory_order_relaxed)) return 0;
  State& s = S();
  int h = 0;
  PS_HIP(hipMemcpyAsync(&h, s.d_err, sizeof(int), hipMemcpyDeviceToHost, st));
  PS_HIP(hipStreamSynchronize(st));
  if (h) {
    hipLaunchKernelGGL(ps_set_i32_k, dim3(1), dim3(1), 0, st, s.d_err, 0);
    PS_HIP(hipStreamSynchronize(st));
    return fail(PS_EBOUNDS, "BoundsError: a kernel accessed an array outside its bounds (the access was suppressed)");
  }
  return 0;
}

extern "C" int ps_device_sync(void) {
  if (int r = require_init()) return r;
  PS_HIP(hipDeviceSynchronize());
  return check_err_flag(S().stream0);
}

// ---------------------------------------------------------------------------
// fields
// ---------------------------------------------------------------------------
// Allocation clears the new memory on the library stream and waits for it; while that stream is
// being captured into a graph the clear would become a graph node and the wait would break the
// capture, so allocating then is refused.
static int refuse_during_capture(const char* what) {
  hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
  PS_HIP(hipStreamIsCapturing(S().stream0, &cs));
  if (cs != hipStreamCaptureStatusNone)
    return fail(PS_EARGUMENT, std::string("cannot ") + what + " while a graph is being captured on the library stream");
  return 0;
}

extern "C" int ps_field_alloc(int ndims, const int64_t dims[3], int dtype, ps_field* out) {
  if (int r = require_init()) return r;
  if (ndims < 1 || ndims > 3) return fail(PS_EARGUMENT, "ps_field_alloc: ndims must be 1, 2 or 3");
  size_t es = elsize(dtype);
  if (!es) return fail(PS_EARGUMENT, "ps_field_alloc: unsupported element type");
  int64_t d[3] = {1, 1, 1};
  for (int i = 0; i < ndims; ++i) {
    if (dims[i] < 0) return fail(PS_EARGUMENT, "ps_field_alloc: negative dimension");
    d[i] = dims[i];
  }
  int64_t align_el = 256 / (int64_t)es;
  int64_t pitch = ((std::max<int64_t>(d[0], 1) + align_el - 1) / align_el) * align_el;
  if (d[1] == 1 && d[2] == 1) pitch = std::max<int64_t>(d[0], 1);  // 1-D fields are dense
  // Rows whose byte stride is a multiple of 4 KiB put the neighbours of a stencil (y +/- 1,
  // z +/- 1) on the same HBM channels; a pad of g_pitch_pad bytes breaks that regularity.
  if (d[1] > 1 && g_pitch_pad > 0 && ((pitch * (int64_t)es) % 4096) == 0) pitch += g_pitch_pad / (int64_t)es;
  int64_t plane = pitch * d[1];
  int64_t bytes = std::max<int64_t>(plane * d[2] * (int64_t)es, 256);
  if (int r = refuse_during_capture("allocate a field")) return r;
  void* p = nullptr;
  hipError_t e = hipMalloc(&p, (size_t)bytes);
  if (e != hipSuccess) return fail(PS_EOOM, "OutOfMemoryError: cannot allocate " + std::to_string(bytes) + " bytes of device memory");
  PS_HIP(hipMemsetAsync(p, 0, (size_t)bytes, S().stream0));
  PS_HIP(hipStreamSynchronize(S().stream0));
  std::memset(out, 0, sizeof *out);
  out->ptr = p;
  out->base = p;
  for (int i = 0; i < 3; ++i) out->dims[i] = d[i];
  out->pitch = pitch;
  out->plane = plane;
  out->dtype = dtype;
  out->ndims = ndims;
  out->alloc_bytes = bytes;
  return 0;
}

extern "C" int ps_cellarray_alloc(int ndims, const int64_t dims[3], int cellndims, const int32_t celldims[4], int64_t blocklength, int dtype,
                                  ps_field* out) {
  if (int r = require_init()) return r;
  if (ndims < 1 || ndims > 3) return fail(PS_EARGUMENT, "ps_cellarray_alloc: ndims must be 1, 2 or 3");
  if (cellndims < 1 || cellndims > 4) return fail(PS_EARGUMENT, "ps_cellarray_alloc: celldims must have 1 to 4 dimensions");
  if (blocklength < 0) return fail(PS_EARGUMENT, "ps_cellarray_alloc: blocklength must be >= 0");
  size_t es = elsize(dtype);
  if (!es) return fail(PS_EARGUMENT, "ps_cellarray_alloc: unsupported element type");
  int64_t d[3] = {1, 1, 1};
  for (int i = 0; i < ndims; ++i) {
    if (dims[i] < 0) return fail(PS_EARGUMENT, "ps_cellarray_alloc: negative dimension");
    d[i] = dims[i];
  }
  int32_t cd[4] = {1, 1, 1, 1};
  int64_t nc = 1;
  for (int i = 0; i < cellndims; ++i) {
    if (celldims[i] < 1) return fail(PS_EARGUMENT, "ps_cellarray_alloc: cell dimensions must be >= 1");
    cd[i] = celldims[i];
    nc *= celldims[i];
  }
  if (nc > 4096) return fail(PS_EARGUMENT, "ps_cellarray_alloc: at most 4096 values per cell are supported");
  const int64_t N = d[0] * d[1] * d[2];
  const int64_t B = blocklength == 0 ? std::max<int64_t>(N, 1) : blocklength;
  const int64_t nblk = (N + B - 1) / B;
  const int64_t bytes = std::max<int64_t>(nblk * B * nc * (int64_t)es, 256);
  if (int r = refuse_during_capture("allocate a field")) return r;
  void* p = nullptr;
  hipError_t e = hipMalloc(&p, (size_t)bytes);
  if (e != hipSuccess) return fail(PS_EOOM, "OutOfMemoryError: cannot allocate " + std::to_string(bytes) + " bytes of device memory");
  PS_HIP(hipMemsetAsync(p, 0, (size_t)bytes, S().stream0));
  PS_HIP(hipStreamSynchronize(S().stream0));
  std::memset(out, 0, sizeof *out);
  out->ptr = p;
  out->base = p;
  for (int i = 0; i < 3; ++i) out->dims[i] = d[i];
  out->pitch = d[0];
  out->plane = d[0] * d[1];
  out->dtype = dtype;
  out->ndims = ndims;
  out->alloc_bytes = bytes;
  out->ncomp = (int32_t)nc;
  for (int i = 0; i < 4; ++i) out->celldims[i] = cd[i];
  out->blocklen = B;
  return 0;
}

extern "C" int ps_field_free(ps_field* f) {
  if (!f || !f->base) return 0;
  if (int r = require_init()) return r;
  PS_HIP(hipFree(f->base));
  f->base = nullptr;
  f->ptr = nullptr;
  return 0;
}

extern "C" int ps_field_view(const ps_field* parent, const int64_t lo[3], const int64_t hi[3], ps_field* out) {
  if (parent->ncomp > 0) return fail(PS_EARGUMENT, "ps_field_view: views of CellArrays are not supported");
  for (int i = 0; i < 3; ++i)
    if (lo[i] < 1 || hi[i] > parent->dims[i] || hi[i] < lo[i] - 1) return fail(PS_EARGUMENT, "ps_field_view: index box outside the parent field");
  *out = *parent;
  size_t es = elsize(parent->dtype);
  out->ptr = (char*)parent->ptr + ((lo[0] - 1) + (lo[1] - 1) * parent->pitch + (lo[2] - 1) * parent->plane) * (int64_t)es;
  for (int i = 0; i < 3; ++i) {
    out->dims[i] = hi[i] - lo[i] + 1;
    out->voff[i] = voffset(*parent, i) + lo[i] - 1;  // a view of a view lies in the same parent
  }
  for (int i = 0; i < 3; ++i) out->pdims[i] = pdim(*parent, i);
  out->base = nullptr;
  out->alloc_bytes = 0;
  return 0;
}

extern "C" int ps_field_h2d(const ps_field* f_in, const void* host, void* stream) {
  if (int r = require_init()) return r;
  const ps_field fd = data_of(*f_in);
  const ps_field* f = &fd;
  size_t es = elsize(f->dtype);
  hipStream_t st = stream_of(stream);
  const char* h = (const char*)host;
  size_t w = (size_t)f->dims[0] * es;
  if (f->plane == f->pitch * f->dims[1]) {
    PS_HIP(hipMemcpy2DAsync(f->ptr, (size_t)f->pitch * es, h, w, w, (size_t)(f->dims[1] * f->dims[2]), hipMemcpyHostToDevice, st));
  } else {
    for (int64_t z = 0; z < f->dims[2]; ++z)
      PS_HIP(hipMemcpy2DAsync((char*)f->ptr + z * f->plane * es, (size_t)f->pitch * es, h + z * w * f->dims[1], w, w, (size_t)f->dims[1],
                              hipMemcpyHostToDevice, st));
  }
  PS_HIP(hipStreamSynchronize(st));
  return 0;
}

extern "C" int ps_field_d2h(const ps_field* f_in, void* host, void* stream) {
  if (int r = require_init()) return r;
  const ps_field fd = data_of(*f_in);
  const ps_field* f = &fd;
  size_t es = elsize(f->dtype);
  hipStream_t st = stream_of(stream);
  char* h = (char*)host;
  size_t w = (size_t)f->dims[0] * es;
  if (f->plane == f->pitch * f->dims[1]) {
    PS_HIP(hipMemcpy2DAsync(h, w, f->ptr, (size_t)f->pitch * es, w, (size_t)(f->dims[1] * f->dims[2]), hipMemcpyDeviceToHost, st));
  } else {
    for (int64_t z = 0; z < f->dims[2]; ++z)
      PS_HIP(hipMemcpy2DAsync(h + z * w * f->dims[1], w, (const char*)f->ptr + z * f->plane * es, (size_t)f->pitch * es, w, (size_t)f->dims[1],
                              hipMemcpyDeviceToHost, st));
  }
  PS_HIP(hipStreamSynchronize(st));
  return 0;
}

static unsigned grid_for(long long n) {
  long long b = (n + 255) / 256;
  return (unsigned)std::max<long long>(1, std::min<long long>(b, 65536));
}

extern "C" int ps_field_fill(const ps_field* f_in, double value, void* stream) {
  if (int r = require_init()) return r;
  const ps_field fd = data_of(*f_in);
  const ps_field* f = &fd;
  hipStream_t st = stream_of(stream);
  long long n = f->dims[0] * f->dims[1] * f->dims[2];
  if (n == 0) return 0;
  dim3 g(grid_for(n)), b(256);
  switch (f->dtype) {
    case PS_F64: hipLaunchKernelGGL(ps_fill_k<double>, g, b, 0, st, (double*)f->ptr, f->dims[0], f->dims[1], f->dims[2], f->pitch, f->plane, value); break;
    case PS_F32: hipLaunchKernelGGL(ps_fill_k<float>, g, b, 0, st, (float*)f->ptr, f->dims[0], f->dims[1], f->dims[2], f->pitch, f->plane, (float)value); break;
    case PS_BF16: {
      float fv = (float)value;
      unsigned u;
      std::memcpy(&u, &fv, 4);
      unsigned short h = (u & 0x7fffffffu) > 0x7f800000u ? (unsigned short)0x7fc0u : (unsigned short)((u + 0x7fffu + ((u >> 16) & 1u)) >> 16);
      hipLaunchKernelGGL(ps_fill_k<unsigned short>, g, b, 0, st, (unsigned short*)f->ptr, f->dims[0], f->dims[1], f->dims[2], f->pitch, f->plane, h);
      break;
    }
    case PS_I32: hipLaunchKernelGGL(ps_fill_k<int>, g, b, 0, st, (int*)f->ptr, f->dims[0], f->dims[1], f->dims[2], f->pitch, f->plane, (int)value); break;
    case PS_I64: hipLaunchKernelGGL(ps_fill_k<long long>, g, b, 0, st, (long long*)f->ptr, f->dims[0], f->dims[1], f->dims[2], f->pitch, f->plane, (long long)value); break;
    case PS_BOOL: hipLaunchKernelGGL(ps_fill_k<unsigned char>, g, b, 0, st, (unsigned char*)f->ptr, f->dims[0], f->dims[1], f->dims[2], f->pitch, f->plane, (unsigned char)(value != 0)); break;
    default: return fail(PS_EARGUMENT, "ps_field_fill: unsupported element type");
  }
  PS_HIP(hipGetLastError());
  return 0;
}

extern "C" int ps_field_copy(const ps_field* dst_in, const ps_field* src_in, void* stream) {
  if (int r = require_init()) return r;
  if ((dst_in->ncomp > 0) != (src_in->ncomp > 0) ||
      (dst_in->ncomp > 0 && (dst_in->ncomp != src_in->ncomp || dst_in->blocklen != src_in->blocklen)))
    return fail(PS_EARGUMENT, "copy!: CellArrays differ in cell size or block length");
  const ps_field dd = data_of(*dst_in), sd = data_of(*src_in);
  const ps_field *dst = &dd, *src = &sd;
  for (int i = 0; i < 3; ++i)
    if (dst->dims[i] != src->dims[i]) return fail(PS_EARGUMENT, "copy!: fields differ in size");
  if (dst->dtype != src->dtype) return fail(PS_EARGUMENT, "copy!: fields differ in element type");
  hipStream_t st = stream_of(stream);
  long long n = dst->dims[0] * dst->dims[1] * dst->dims[2];
  if (n == 0) return 0;
  dim3 g(grid_for(n)), b(256);
  size_t es = elsize(dst->dtype);
#define PS_COPY(T) hipLaunchKernelGGL(ps_copy_k<T>, g, b, 0, st, (T*)dst->ptr, (const T*)src->ptr, dst->dims[0], dst->dims[1], dst->dims[2], dst->pitch, dst->plane, src->pitch, src->plane)
  if (es == 8) PS_COPY(long long);
  else if (es == 4) PS_COPY(int);
  else if (es == 2) PS_COPY(unsigned short);
  else PS_COPY(unsigned char);
#undef PS_COPY
  PS_HIP(hipGetLastError());
  return 0;
}

extern "C" int ps_field_reduce(const ps_field* f_in, int op, double* out, void* stream) {
  if (int r = require_init()) return r;
  const ps_field fd = data_of(*f_in);
  const ps_field* f = &fd;
  if (op < 0 || op > 4) return fail(PS_EARGUMENT, "ps_field_reduce: op must be 0..4");
  State& s = S();
  hipStream_t st = stream_of(stream);
  long long n = f->dims[0] * f->dims[1] * f->dims[2];
  int nb = (int)std::min<long long>(std::max<long long>((n + 255) / 256, 1), kReduceParts);
  switch (f->dtype) {
    case PS_F64: hipLaunchKernelGGL(ps_reduce1_k<double>, dim3(nb), dim3(256), 0, st, (const double*)f->ptr, f->dims[0], f->dims[1], f->dims[2], f->pitch, f->plane, s.d_part); break;
    case PS_F32: hipLaunchKernelGGL(ps_reduce1_k<float>, dim3(nb), dim3(256), 0, st, (const float*)f->ptr, f->dims[0], f->dims[1], f->dims[2], f->pitch, f->plane, s.d_part); break;
    case PS_BF16: hipLaunchKernelGGL(ps_reduce1_k<unsigned short>, dim3(nb), dim3(256), 0, st, (const unsigned short*)f->ptr, f->dims[0], f->dims[1], f->dims[2], f->pitch, f->plane, s.d_part); break;
    default: return fail(PS_EARGUMENT, "ps_field_reduce: floating-point fields only");
  }
  hipLaunchKernelGGL(ps_reduce2_k, dim3(1), dim3(256), 0, st, s.d_part, nb, s.d_out);
  PS_HIP(hipGetLastError());
  double h[4];
  PS_HIP(hipMemcpyAsync(h, s.d_out, 4 * sizeof(double), hipMemcpyDeviceToHost, st));
  PS_HIP(hipStreamSynchronize(st));
  if (op == 0) *out = h[0];
  else if (op == 1) *out = h[1];
  else if (op == 2) *out = h[2];
  else if (op == 3) *out = n ? h[1] / (double)n : 0.0;
  else *out = std::sqrt(h[3]);
  return 0;
}

// The variant of ps_field_reduce for callers that cannot allocate: up to kNoallocFields fields per
// call, asynchronous on `stream`, results into the caller's device memory out[0..nfields-1]. The
// partials go to the library's static scratch array of the current device; nothing is allocated,
// copied to the host or waited for, so the call can be captured into a graph. Every partial that
// stage 2 reads is written by stage 1 of the same call, so nothing needs clearing between calls.
// Calls that overlap on one GPU share the scratch array and race.
extern "C" int ps_field_reduce_noalloc(const ps_field* fields, int nfields, int op, double* out, void* stream) {
  if (int r = require_init()) return r;
  if (nfields < 1 || nfields > kNoallocFields)
    return fail(PS_EARGUMENT, "ps_field_reduce_noalloc: nfields must be 1.." + std::to_string(kNoallocFields) + " (obtained " +
                                  std::to_string(nfields) + "); the scratch array holds the partials of at most " +
                                  std::to_string(kNoallocFields) + " fields");
  if (op < 0 || op > 4) return fail(PS_EARGUMENT, "ps_field_reduce_noalloc: op must be 0..4");
  if (!fields || !out) return fail(PS_EARGUMENT, "ps_field_reduce_noalloc: fields and out must not be NULL");
  State& s = S();
  auto it = s.scratch_ptr.find(s.device);
  if (it == s.scratch_ptr.end()) return fail(PS_EINTERNAL, "ps_field_reduce_noalloc: no scratch address for this device");
  double* part = it->second;
  hipStream_t st = stream_of(stream);
  ps_red_meta meta{};
  for (int i = 0; i < nfields; ++i) {
    const ps_field fd = data_of(fields[i]);
    const long long n = fd.dims[0] * fd.dims[1] * fd.dims[2];
    if (n > 0 && !fd.ptr) return fail(PS_EARGUMENT, "ps_field_reduce_noalloc: field " + std::to_string(i + 1) + " has no device memory");
    if (fd.dtype != PS_F64 && fd.dtype != PS_F32 && fd.dtype != PS_BF16)
      return fail(PS_EARGUMENT, "ps_field_reduce_noalloc: floating-point fields only");
    meta.nparts[i] = (int)std::min<long long>(std::max<long long>((n + 255) / 256, 1), kReduceParts);
    meta.n[i] = n;
  }
  for (int i = 0; i < nfields; ++i) {
    const ps_field fd = data_of(fields[i]);
    double* pf = part + (size_t)i * kReduceParts * 4;
    const dim3 g(meta.nparts[i]), b(256);
    switch (fd.dtype) {
      case PS_F64: hipLaunchKernelGGL(ps_reduce1_k<double>, g, b, 0, st, (const double*)fd.ptr, fd.dims[0], fd.dims[1], fd.dims[2], fd.pitch, fd.plane, pf); break;
      case PS_F32: hipLaunchKernelGGL(ps_reduce1_k<float>, g, b, 0, st, (const float*)fd.ptr, fd.dims[0], fd.dims[1], fd.dims[2], fd.pitch, fd.plane, pf); break;
      default: hipLaunchKernelGGL(ps_reduce1_k<unsigned short>, g, b, 0, st, (const unsigned short*)fd.ptr, fd.dims[0], fd.dims[1], fd.dims[2], fd.pitch, fd.plane, pf); break;
    }
  }
  hipLaunchKernelGGL(ps_reduce2_op_k, dim3(nfields), dim3(256), 0, st, (const double*)part, meta, op, out);
  PS_HIP(hipGetLastError());
  return 0;
}

// ---------------------------------------------------------------------------
// streams and events
// ---------------------------------------------------------------------------
extern "C" int ps_stream_get(int id, int priority, void** out) {
  if (int r = require_init()) return r;
  if (id < 1) return fail(PS_EARGUMENT, "stream ids start at 1");
  State& s = S();
  std::lock_guard<std::mutex> lk(s.mu);
  auto& pool = priority ? s.prio_streams : s.streams;
  while ((int)pool.size() < id) {
    hipStream_t st;
    PS_HIP(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
    pool.push_back(st);
  }
  *out = (void*)pool[id - 1];
  return 0;
}

extern "C" int ps_stream_sync(void* stream) {
  if (int r = require_init()) return r;
  hipStream_t st = stream_of(stream);
  PS_HIP(hipStreamSynchronize(st));
  return check_err_flag(st);
}

extern "C" int ps_event_create(void** ev) {
  if (int r = require_init()) return r;
  hipEvent_t e;
  PS_HIP(hipEventCreate(&e));
  *ev = (void*)e;
  return 0;
}
extern "C" int ps_event_record(void* ev, void* stream) {
  PS_HIP(hipEventRecord((hipEvent_t)ev, stream_of(stream)));
  return 0;
}
extern "C" int ps_event_elapsed_ms(void* ev0, void* ev1, float* ms) {
  PS_HIP(hipEventSynchronize((hipEvent_t)ev1));
  PS_HIP(hipEventElapsedTime(ms, (hipEvent_t)ev0, (hipEvent_t)ev1));
  return 0;
}
extern "C" int ps_event_destroy(void* ev) {
  PS_HIP(hipEventDestroy((hipEvent_t)ev));
  return 0;
}

extern "C" int ps_set_cache_dir(const char* dir) {
  S().cache_dir = dir ? dir : "";
  return 0;
}

// ---------------------------------------------------------------------------
// graphs: capture the asynchronous launches issued on a stream, replay them as one launch
// ---------------------------------------------------------------------------
extern "C" int ps_graph_begin(void* stream) {
  if (int r = require_init()) return r;
  PS_HIP(hipStreamBeginCapture(stream_of(stream), hipStreamCaptureModeRelaxed));
  return 0;
}

extern "C" int ps_graph_end(void* stream, void** exec) {
  if (int r = require_init()) return r;
  hipGraph_t g = nullptr;
  PS_HIP(hipStreamEndCapture(stream_of(stream), &g));
  hipGraphExec_t x = nullptr;
  hipError_t e = hipGraphInstantiate(&x, g, nullptr, nullptr, 0);
  (void)hipGraphDestroy(g);
  if (e != hipSuccess) return fail(PS_EHIP, std::string("hipGraphInstantiate: ") + hipGetErrorString(e));
  *exec = (void*)x;
  return 0;
}

extern "C" int ps_graph_launch(void* exec, void* stream) {
  if (int r = require_init()) return r;
  PS_HIP(hipGraphLaunch((hipGraphExec_t)exec, stream_of(stream)));
  return 0;
}

extern "C" int ps_graph_destroy(void* exec) {
  PS_HIP(hipGraphExecDestroy((hipGraphExec_t)exec));
  return 0;
}

// ---------------------------------------------------------------------------
// geometry helpers
// ---------------------------------------------------------------------------
extern "C" int ps_compute_nthreads(const int64_t maxsize[3], int nthreads_x_max, int nthreads_max, int flatdim, int32_t out[3]) {
  int64_t nx = std::min<int64_t>(nthreads_x_max, flatdim == 1 ? 1 : maxsize[0]);
  if (nx < 1) nx = 1;
  int64_t ny = std::min<int64_t>((nthreads_max + nx - 1) / nx, flatdim == 2 ? 1 : maxsize[1]);
  if (ny < 1) ny = 1;
  int64_t nz = std::min<int64_t>((nthreads_max + nx * ny - 1) / (nx * ny), flatdim == 3 ? 1 : maxsize[2]);
  if (nz < 1) nz = 1;
  out[0] = (int32_t)nx;
  out[1] = (int32_t)ny;
  out[2] = (int32_t)nz;
  return 0;
}

extern "C" int ps_compute_nblocks(const int64_t maxsize[3], const int32_t nthreads[3], int32_t out[3]) {
  for (int i = 0; i < 3; ++i) {
    if (nthreads[i] < 1) return fail(PS_EARGUMENT, "nthreads must be positive");
    out[i] = (int32_t)((maxsize[i] + nthreads[i] - 1) / nthreads[i]);
  }
  return 0;
}

static int validate_hc(const int64_t r[3][2], const int64_t bw[3]) {
  int64_t ms[3];
  for (int i = 0; i < 3; ++i) {
    if (r[i][0] != 1)
      return fail(PS_EARGUMENT, "unsupported call to `@hide_communication boundary_width block`: automatic ranges determination for @hide_communication is not supported for ranges with start != 1. Use the explicit version `@hide_communication ranges_outer ranges_inner block` instead.");
    ms[i] = r[i][1] - r[i][0] + 1;
  }
  for (int i = 0; i < 3; ++i)
    if (ms[i] <= 0) return fail(PS_EINCOHERENT, "incoherent arguments in @hide_communication: the following must be true: all(length.(ranges) .> 0) ");
  for (int i = 0; i < 3; ++i)
    if (!(ms[i] > 2 * bw[i])) return fail(PS_EINCOHERENT, "incoherent arguments in @hide_communication: the following must be true: all(length.(ranges) .> 2 .* boundary_width) ");
  for (int i = 0; i < 3; ++i)
    if (ms[i] > 1 && bw[i] == 0) return fail(PS_EINCOHERENT, "incoherent arguments in @hide_communication: the following must be false: any((length.(ranges) .> 1) .& (boundary_width .== 0)) ");
  if (ms[0] == 1 && ms[1] == 1 && ms[2] == 1) return fail(PS_EINCOHERENT, "incoherent arguments in @hide_communication: the following must be false: all(length.(ranges) .== 1)");
  return 0;
}

extern "C" int ps_ranges_outer(const int64_t r[3][2], const int64_t bw[3], int64_t out[6][3][2], int* nout) {
  if (int e = validate_hc(r, bw)) return e;
  int64_t ms[3] = {r[0][1], r[1][1], r[2][1]};
  int64_t boxes[6][3][2];
  int n = 0;
  auto add = [&](int64_t a0, int64_t a1, int64_t b0, int64_t b1, int64_t c0, int64_t c1) {
    boxes[n][0][0] = a0; boxes[n][0][1] = a1;
    boxes[n][1][0] = b0; boxes[n][1][1] = b1;
    boxes[n][2][0] = c0; boxes[n][2][1] = c1;
    ++n;
  };
  if (ms[2] > 1) {
    add(1, bw[0], bw[1] + 1, ms[1] - bw[1], bw[2] + 1, ms[2] - bw[2]);
    add(1, ms[0], 1, ms[1], 1, bw[2]);
    add(1, ms[0], 1, ms[1], ms[2] - bw[2] + 1, ms[2]);
    add(ms[0] - bw[0] + 1, ms[0], bw[1] + 1, ms[1] - bw[1], bw[2] + 1, ms[2] - bw[2]);
    add(1, ms[0], 1, bw[1], bw[2] + 1, ms[2] - bw[2]);
    add(1, ms[0], ms[1] - bw[1] + 1, ms[1], bw[2] + 1, ms[2] - bw[2]);
  } else if (ms[1] > 1) {
    add(ms[0] - bw[0] + 1, ms[0], bw[1] + 1, ms[1] - bw[1], 1, 1);
    add(1, ms[0], 1, bw[1], 1, 1);
    add(1, ms[0], ms[1] - bw[1] + 1, ms[1], 1, 1);
    add(1, bw[0], bw[1] + 1, ms[1] - bw[1], 1, 1);
  } else {
    add(ms[0] - bw[0] + 1, ms[0], 1, 1, 1, 1);
    add(1, bw[0], 1, 1, 1, 1);
  }
  int k = 0;
  for (int i = 0; i < n; ++i) {
    bool nonempty = true;
    for (int d = 0; d < 3; ++d)
      if (boxes[i][d][1] - boxes[i][d][0] + 1 == 0) nonempty = false;
    if (!nonempty) continue;
    std::memcpy(out[k], boxes[i], sizeof boxes[i]);
    ++k;
  }
  *nout = k;
  return 0;
}

extern "C" int ps_ranges_inner(const int64_t r[3][2], const int64_t bw[3], int64_t out[3][2]) {
  if (int e = validate_hc(r, bw)) return e;
  int64_t ms[3] = {r[0][1], r[1][1], r[2][1]};
  for (int d = 0; d < 3; ++d) {
    out[d][0] = bw[d] + 1;
    out[d][1] = ms[d] - bw[d];
  }
  if (ms[2] <= 1) out[2][0] = out[2][1] = 1;
  if (ms[1] <= 1) out[1][0] = out[1][1] = 1;
  return 0;
}

// ---------------------------------------------------------------------------
// kernels
// ---------------------------------------------------------------------------
extern "C" int ps_kernel_define(const char* text, const char* opts_json, ps_kernel* out) {
  PS_GUARD_BEGIN
  auto ko = new KernelObj();
  try {
    ko->K = define_kernel(text ? text : "", json_parse(opts_json ? opts_json : ""));
  } catch (...) {
    delete ko;
    throw;
  }
  *out = (ps_kernel)ko;
  return 0;
  PS_GUARD_END
}

extern "C" int ps_kernel_derivative(ps_kernel primal, const char* spec_json, ps_kernel* out) {
  PS_GUARD_BEGIN
  KernelObj* p = (KernelObj*)primal;
  if (!p || !out) throw PsError(PS_EARGUMENT, "ps_kernel_derivative: null kernel");
  auto ko = new KernelObj();
  try {
    ko->K = derivative_kernel(p->K, json_parse(spec_json ? spec_json : ""));
  } catch (...) {
    delete ko;
    throw;
  }
  *out = (ps_kernel)ko;
  return 0;
  PS_GUARD_END
}

extern "C" int ps_kernel_free(ps_kernel k) {
  KernelObj* ko = (KernelObj*)k;
  if (!ko) return 0;
  for (auto& v : ko->variants)
    if (v.second->mod) (void)hipModuleUnload(v.second->mod);
  delete ko;
  return 0;
}

extern "C" int ps_kernel_info(ps_kernel k, char* out, size_t cap) {
  PS_GUARD_BEGIN
  std::string s = kernel_info_json(((KernelObj*)k)->K);
  if (s.size() + 1 > cap) return fail(PS_EARGUMENT, "ps_kernel_info: buffer too small (need " + std::to_string(s.size() + 1) + ")");
  std::memcpy(out, s.c_str(), s.size() + 1);
  return 0;
  PS_GUARD_END
}

// Host evaluation of a hoisted scalar subexpression (Emitted::UProg) in the arithmetic the device
// would use for it: every operation in its result type (float or double) after converting both
// operands to that type, integers truncated to ps_int, no contraction. IEEE +, -, *, / are
// correctly rounded on both sides, so the value is the one each cell would have computed.
namespace {
struct UVal {
  int ty;  // 0 ps_int, 1 float, 2 double
  long long i;
  float f;
  double d;
};
UVal uval_to(const UVal& v, int ty) {
  if (v.ty == ty) return v;
  UVal r{ty, 0, 0.0f, 0.0};
  if (ty == 1 && v.ty == 0) r.f = (float)v.i;
  else if (ty == 2 && v.ty == 0) r.d = (double)v.i;
  else if (ty == 2 && v.ty == 1) r.d = (double)v.f;
  else throw PsError(PS_EINTERNAL, "hoisted subexpression narrows a value");
  return r;
}
template <class T>
T uarith(int op, T a, T b) {
#pragma clang fp contract(off)
  using U = Emitted::UOp;
  switch (op) {
    case U::Add: return a + b;
    case U::Sub: return a - b;
    case U::Mul: return a * b;
    case U::Div: return a / b;
  }
  throw PsError(PS_EINTERNAL, "hoisted subexpression: bad operation");
}
template <class T>
T uunary(int op, T a) {
#pragma clang fp contract(off)
  using U = Emitted::UOp;
  switch (op) {
    case U::Neg: return -a;
    case U::Pow0: return (T)1;
    case U::Pow2: return a * a;
    case U::Pow3: return a * a * a;
    case U::PowM1: return (T)1 / a;
    case U::PowM2: {
      const T i = (T)1 / a;
      return i * i;
    }
  }
  throw PsError(PS_EINTERNAL, "hoisted subexpression: bad operation");
}
UVal eval_uprog(const Emitted::UProg& p, const ps_arg* args) {
  using U = Emitted::UOp;
  std::vector<UVal> st;
  for (const U& u : p.ops) {
    if (u.op == U::Arg || u.op == U::Lit) {
      UVal v{u.ty, 0, 0.0f, 0.0};
      const double fv = u.op == U::Arg ? args[u.arg].f : u.lit;
      const long long iv = u.op == U::Arg ? args[u.arg].i : u.ilit;
      if (u.ty == 2) v.d = fv;
      else if (u.ty == 1) v.f = (float)fv;
      else v.i = p.idx64 ? iv : (long long)(int32_t)iv;
      st.push_back(v);
    } else if (u.op >= U::Add && u.op <= U::Div) {
      if (st.size() < 2) throw PsError(PS_EINTERNAL, "malformed hoisted subexpression");
      const UVal b = uval_to(st.back(), u.ty);
      st.pop_back();
      const UVal a = uval_to(st.back(), u.ty);
      st.pop_back();
      UVal r{u.ty, 0, 0.0f, 0.0};
      if (u.ty == 1) r.f = uarith<float>(u.op, a.f, b.f);
      else if (u.ty == 2) r.d = uarith<double>(u.op, a.d, b.d);
      else throw PsError(PS_EINTERNAL, "hoisted subexpression: integer arithmetic");
      st.push_back(r);
    } else {
      if (st.empty()) throw PsError(PS_EINTERNAL, "malformed hoisted subexpression");
      UVal& x = st.back();
      if (x.ty == 1) x.f = uunary<float>(u.op, x.f);
      else if (x.ty == 2) x.d = uunary<double>(u.op, x.d);
      else throw PsError(PS_EINTERNAL, "hoisted subexpression: integer arithmetic");
    }
  }
  if (st.size() != 1 || st.back().ty != p.ty) throw PsError(PS_EINTERNAL, "malformed hoisted subexpression");
  return st.back();
}
double eval_uprog_f64(const Emitted::UProg& p, const ps_arg* args) { return eval_uprog(p, args).d; }
float eval_uprog_f32(const Emitted::UProg& p, const ps_arg* args) { return eval_uprog(p, args).f; }
}  // namespace

// Division mode of every argument (see Variant::divmode): float scalars whose magnitude the
// reciprocal division handles exactly are 'p' / 'n' by sign, other float scalars 'x'.
static std::string divmode_of(const ps_arg* args, int nargs) {
  std::string m(nargs, '-');
  for (int i = 0; i < nargs; ++i) {
    if (args[i].kind != PS_ARG_SCALAR) continue;
    if (args[i].dtype == PS_F64) m[i] = rcp_safe64(args[i].f) ? (args[i].f > 0 ? 'p' : 'n') : 'x';
    else if (args[i].dtype == PS_F32) {
      float b = (float)args[i].f;
      m[i] = rcp_safe32(b) ? (b > 0 ? 'p' : 'n') : 'x';
    }
  }
  return m;
}

static Variant variant_from_flags(const Kernel& K, int flags) {
  Variant v;
  v.memopt = (flags & 1) ? true : false;
  if (v.memopt && !K.mm.enabled) throw PsError(PS_EINCOHERENT, "the kernel was not defined with memopt=true");
  v.checked = (flags & 2) != 0;
  v.idx64 = (flags & 4) != 0;
  v.noalias = (flags & 8) == 0;
  v.maxthreads = (flags & 16) ? 1024 : 256;
  if (flags & 32) {  // LDS-tiled memopt variant for the block shape in bits 8..19 / 20..31
    v.lds = true;
    v.bx = (flags >> 8) & 0xfff;
    v.by = (flags >> 20) & 0xfff;
    if (!v.memopt || v.bx < 1 || v.by < 1) throw PsError(PS_EARGUMENT, "LDS variants need memopt and a block shape");
  }
  v.nt = K.opt.nt_store == 1 || (K.opt.nt_store == -1 && v.lds && v.bx >= 256);
  v.off32 = (flags & 64) != 0 && !v.memopt && !v.checked;  // 32-bit per-thread byte offsets
  v.u24 = v.off32 && (flags & 128) != 0;                      // 24-bit cell-offset multiplies
  // groups: x cells per thread in bits 8..11, z cells in bits 12..15
  if (v.off32 && !v.lds && ((flags >> 8) & 0xf) > 1 && groupable_body(K)) v.cx = (flags >> 8) & 0xf;
  if (v.off32 && !v.lds && ((flags >> 12) & 0xf) > 1 && groupable_body(K, true)) v.cz = (flags >> 12) & 0xf;
  return v;
}

extern "C" int ps_kernel_source(ps_kernel k, const ps_arg* args, int nargs, int flags, char* out, size_t cap) {
  PS_GUARD_BEGIN
  KernelObj* ko = (KernelObj*)k;
  Variant v = variant_from_flags(ko->K, flags);
  v.divmode = divmode_of(args, nargs);
  const Kernel K0 = ko->K.pgt_terms.empty() ? ko->K : specialize_padding(ko->K, std::string(ko->K.pgt_terms.size(), '0'));
  if (!ko->K.pgt_terms.empty()) v.padbits = std::string(ko->K.pgt_terms.size(), '0');
  Emitted em = lower_kernel(K0, arg_kinds(args, nargs), v);
  if (em.source.size() + 1 > cap) return fail(PS_EARGUMENT, "ps_kernel_source: buffer too small (need " + std::to_string(em.source.size() + 1) + ")");
  std::memcpy(out, em.source.c_str(), em.source.size() + 1);
  return 0;
  PS_GUARD_END
}

extern "C" int ps_kernel_precompile(ps_kernel k, const ps_arg* args, int nargs, int flags) {
  PS_GUARD_BEGIN
  KernelObj* ko = (KernelObj*)k;
  Variant v = variant_from_flags(ko->K, flags);
  v.divmode = divmode_of(args, nargs);
  if (!ko->K.pgt_terms.empty()) {
    v.padbits = std::string(ko->K.pgt_terms.size(), '0');
    get_module(ko, specialize_padding(ko->K, v.padbits), arg_kinds(args, nargs), v, false);
  } else {
    get_module(ko, ko->K, arg_kinds(args, nargs), v, false);
  }
  return 0;
  PS_GUARD_END
}

namespace {

struct Plan {
  bool memopt = false;
  int64_t ranges[3][2];
  int32_t nthreads[3];
  int32_t nblocks[3];  // the reference's geometry (what geometry() reports)
  // the launched block and grid: the reference's geometry, or for automatic geometry of plain
  // kernels the x-group block shape and the grid padded for the XCDs
  int32_t lthreads[3];
  int32_t lblocks[3];
  Variant v;
};

static const char* kErrAutoRanges =
    "automatic detection of ranges not possible in @parallel <kernelcall>: some kernel arguments are neither arrays nor scalars nor any other bitstypes nor (named) tuple containing any of the former. Specify ranges or nthreads and nblocks manually.";
static const char* kErrNbIdx =
    "the number of parallel indices of the kernel does not match the number of dimensions of the kernel arguments (maximal number of dimensions of array arguments); specify the ranges explicitly or use @parallel_indices with the matching number of indices.";

Plan plan_launch(const Kernel& K, const ps_arg* args, int nargs, const ps_launch* L, std::vector<ArgKind>& kinds,
                 const std::string& padbits = "") {
  (void)kErrAutoRanges;
  Plan P;
  kinds = arg_kinds(args, nargs);
  const bool pad = K.opt.padding;
  if (pad)
    for (int i = 0; i < nargs; ++i)
      if (args[i].kind == PS_ARG_FIELD && args[i].field.ncomp > 0)
        throw PsError(PS_EARGUMENT, "padding=true is not supported with CellArray arguments (argument " + std::to_string(i + 1) + ")");
  if (nargs != (int)K.def.params.size())
    throw PsError(PS_EARGUMENT, "MethodError: kernel " + K.def.name + " takes " + std::to_string(K.def.params.size()) + " arguments (" +
                                    std::to_string(nargs) + " given)");
  const int64_t shmem = L ? L->shmem : 0;
  if (shmem < 0) throw PsError(PS_EARGUMENT, "@parallel <kernelcall>: shmem must be a non-negative number of bytes");
  if (!K.shared_arrays.empty() && shmem == 0)
    throw PsError(PS_EARGUMENT, "@parallel <kernelcall>: kernel " + K.def.name +
                                    " creates @sharedMem arrays; the amount of shared memory needs to be specified when launching the kernel (keyword argument `shmem`)");
  bool memopt = K.mm.enabled;
  if (L && L->memopt == 1 && !memopt) throw PsError(PS_EINCOHERENT, "@parallel memopt=true: the kernel was not defined with memopt=true");
  if (L && L->memopt == 0) memopt = false;
  // memopt kernels load branch-free from a clamped index (element 0 of the array stands in for
  // an access outside it), which needs every array to hold at least one element
  for (int i = 0; i < nargs && memopt; ++i)
    if (args[i].kind == PS_ARG_FIELD && args[i].field.dims[0] * args[i].field.dims[1] * args[i].field.dims[2] == 0) memopt = false;
  for (int i = 0; i < nargs && memopt; ++i)
    if (args[i].kind == PS_ARG_FIELD && args[i].field.ncomp > 0)
      throw PsError(PS_EARGUMENT, "@parallel <kernelcall>: memopt=true is not supported with CellArray arguments (argument " + std::to_string(i + 1) + ")");
  P.memopt = memopt;
  int64_t maxsize[3] = {1, 1, 1};
  int maxnd = 1;
  for (int i = 0; i < nargs; ++i) {  // maxsize: of the parents for padding=true (parallel.jl:786)
    if (args[i].kind != PS_ARG_FIELD) continue;
    for (int d = 0; d < 3; ++d) maxsize[d] = std::max<int64_t>(maxsize[d], pad ? pdim(args[i].field, d) : args[i].field.dims[d]);
    maxnd = std::max(maxnd, (int)args[i].field.ndims);
  }
  bool user_ranges = L && L->has_ranges;
  if (user_ranges) {
    for (int d = 0; d < 3; ++d) {
      P.ranges[d][0] = L->ranges[d][0];
      P.ranges[d][1] = L->ranges[d][1];
    }
  } else {
    if (maxnd != K.nidx) throw PsError(PS_EARGUMENT, kErrNbIdx);
    for (int d = 0; d < 3; ++d) {
      P.ranges[d][0] = 1;
      P.ranges[d][1] = maxsize[d];
    }
  }
  int64_t len[3];
  for (int d = 0; d < 3; ++d) {
    len[d] = P.ranges[d][1] - P.ranges[d][0] + 1;
    if (len[d] < 0) throw PsError(PS_EARGUMENT, "ranges must not be decreasing");
  }
  bool user_geom = L && (L->nthreads[0] || L->nblocks[0]);
  if (memopt) {
    const MemoptMeta& mm = K.mm;
    int Ld = mm.loopdim - 1;
    int64_t loopsizes[3] = {1, 1, 1};
    loopsizes[Ld] = mm.loopsize;
    // block shape: 128 x 4 threads, or 256 x 4 for 3-D streams whose x range is a multiple of
    // 256 of at least 512 cells (wider HBM rows per wave; measured fastest at 512^3), unless the
    // kernel sets nthreads_x_max_memopt / nthreads_max_memopt; then the x width is halved where
    // that pads the x range to fewer threads
    int xmax = K.opt.nthreads_x_max_memopt, tmax = K.opt.nthreads_max_memopt;
    const bool autoshape = xmax == 0 && tmax == 0;
    if (xmax == 0) xmax = 128;
    if (tmax == 0) tmax = 512;
    if (autoshape && mm.loopdim == 3 && len[0] >= 512 && len[0] % 256 == 0) {
      xmax = 256;
      tmax = 1024;
    }
    if (xmax >= 128 && (len[0] + 63) / 64 * 64 < (len[0] + xmax - 1) / xmax * xmax) xmax = 64;
    if (K.opt.kind == "parallel" && !user_ranges) {
      // get_ranges_memopt: pad to multiples of nthreads (src/parallel.jl:560-571)
      int32_t nt[3];
      ps_compute_nthreads(len, xmax, tmax, mm.loopdim, nt);
      for (int d = 0; d < 3; ++d) {
        int64_t rest = len[d] % nt[d];
        if (rest) len[d] += nt[d] - rest;
        P.ranges[d][1] = P.ranges[d][0] + len[d] - 1;
      }
    }
    int64_t ms[3];
    for (int d = 0; d < 3; ++d) ms[d] = (len[d] + loopsizes[d] - 1) / loopsizes[d];
    if (user_geom) {
      for (int d = 0; d < 3; ++d) {
        P.nthreads[d] = L->nthreads[d] ? L->nthreads[d] : 1;
        P.nblocks[d] = L->nblocks[d] ? L->nblocks[d] : (int32_t)((ms[d] + P.nthreads[d] - 1) / P.nthreads[d]);
      }
    } else {
      ps_compute_nthreads(ms, xmax, tmax, mm.loopdim, P.nthreads);
      // sizes that the tuned block does not divide get the reference's AMDGPU block limits
      // (64 threads along x, 256 per block: src/shared.jl:29, src/ParallelKernel/shared.jl:33)
      if (ms[0] % P.nthreads[0] || ms[1] % P.nthreads[1] || ms[2] % P.nthreads[2])
        ps_compute_nthreads(ms, K.opt.nthreads_x_max, 256, mm.loopdim, P.nthreads);
      // compute_nthreads_memopt checks (src/parallel.jl:548-558)
      for (auto& A : mm.optvars) {
        const auto& sr = mm.stencilranges.at(A);
        int64_t he[3];
        for (int d = 0; d < 3; ++d) he[d] = (d == Ld) ? 0 : (sr[d].second - sr[d].first);
        int64_t p1 = (int64_t)P.nthreads[0] * P.nthreads[1] * P.nthreads[2];
        int64_t p2 = (P.nthreads[0] + he[0]) * (P.nthreads[1] + he[1]) * (P.nthreads[2] + he[2]);
        if (2 * p1 < p2)
          throw PsError(PS_EARGUMENT, "@parallel <kernelcall>: the automatic determination of nthreads is not possible for this case. Please specify `nthreads` and `nblocks`.");
      }
      for (int d = 0; d < 3; ++d)
        if (ms[d] % P.nthreads[d] != 0)
          throw PsError(PS_EARGUMENT, "@parallel <kernelcall>: memopt optimization not possible for the given maximum array size in the kernel arguments (the maximum array size must be dividable without rest by the number of threads per block)");
      ps_compute_nblocks(ms, P.nthreads, P.nblocks);
    }
  } else {
    if (user_geom) {
      for (int d = 0; d < 3; ++d) {
        P.nthreads[d] = L->nthreads[d] ? L->nthreads[d] : 1;
        P.nblocks[d] = L->nblocks[d] ? L->nblocks[d] : (int32_t)((len[d] + P.nthreads[d] - 1) / P.nthreads[d]);
      }
    } else {
      ps_compute_nthreads(len, K.opt.nthreads_x_max, K.opt.nthreads_max, 0, P.nthreads);
      ps_compute_nblocks(len, P.nthreads, P.nblocks);
    }
  }
  for (int d = 0; d < 3; ++d) {
    if (P.nthreads[d] < 1 || P.nblocks[d] < 0) throw PsError(PS_EARGUMENT, "invalid nthreads / nblocks");
    P.lthreads[d] = P.nthreads[d];
    P.lblocks[d] = P.nblocks[d];
  }
  if ((int64_t)P.nthreads[0] * P.nthreads[1] * P.nthreads[2] > 1024) throw PsError(PS_EARGUMENT, "at most 1024 threads per block are supported");
  if (P.nblocks[1] > 65535 || P.nblocks[2] > 65535) throw PsError(PS_EARGUMENT, "nblocks in y and z must not exceed 65535");
  // variant
  Variant v;
  v.memopt = memopt;
  if (memopt && K.opt.lds && !K.mm.shmem_optvars.empty() && K.mm.loopdim == 3 && P.nthreads[2] == 1) {
    v.lds = true;
    v.bx = P.nthreads[0];
    v.by = P.nthreads[1];
  }
  v.maxthreads = ((int64_t)P.nthreads[0] * P.nthreads[1] * P.nthreads[2] > 256) ? 1024 : 256;
  v.nt = K.opt.nt_store == 1 || (K.opt.nt_store == -1 && v.lds && v.bx >= 256);
  bool big = false;
  for (int i = 0; i < nargs; ++i) {
    if (args[i].kind == PS_ARG_FIELD) {
      const ps_field& f = args[i].field;
      if (extent_bytes(f) / (int64_t)elsize(f.dtype) + f.pitch > (int64_t)INT32_MAX / 2) big = true;
    } else if (args[i].dtype == PS_I32 || args[i].dtype == PS_I64) {
      if (args[i].i > INT32_MAX / 2 || args[i].i < INT32_MIN / 2) big = true;
    }
  }
  for (int d = 0; d < 3; ++d)
    if (P.ranges[d][1] > INT32_MAX / 2 || P.ranges[d][0] < INT32_MIN / 2) big = true;
  v.idx64 = big || K.needs_idx64;
  v.divmode = divmode_of(args, nargs);
  v.padbits = padbits;
  // aliasing: a written field that overlaps another field argument (padding=true kernels reach the
  // whole parent of every argument)
  for (int i = 0; i < nargs && v.noalias; ++i) {
    if (args[i].kind != PS_ARG_FIELD) continue;
    for (int j = i + 1; j < nargs; ++j) {
      if (args[j].kind != PS_ARG_FIELD) continue;
      const ps_field a = pad ? parent_of(args[i].field) : args[i].field, b = pad ? parent_of(args[j].field) : args[j].field;
      const char* a0 = (const char*)a.ptr;
      const char* a1 = a0 + extent_bytes(a);
      const char* b0 = (const char*)b.ptr;
      const char* b1 = b0 + extent_bytes(b);
      if (a0 < b1 && b0 < a1) v.noalias = false;
    }
  }
  // bounds proof over the index box the threads can reach
  LaunchShape sh;
  sh.kinds = kinds;
  sh.dims.resize(nargs);
  sh.ivals.resize(nargs);
  if (pad) {
    sh.vlo.resize(nargs);
    sh.vhi.resize(nargs);
  }
  for (int i = 0; i < nargs; ++i) {
    if (args[i].kind == PS_ARG_FIELD) {
      for (int d = 0; d < 3; ++d) {
        sh.dims[i][d] = pad ? pdim(args[i].field, d) : args[i].field.dims[d];
        if (pad) {
          sh.vlo[i][d] = voffset(args[i].field, d) + 1;
          sh.vhi[i][d] = voffset(args[i].field, d) + args[i].field.dims[d];
        }
      }
    } else {
      sh.ivals[i] = args[i].i;
    }
  }
  for (int d = 0; d < 3; ++d) {
    sh.lo[d] = P.ranges[d][0];
    int64_t reach = (int64_t)P.nthreads[d] * P.nblocks[d];
    if (memopt && d == K.mm.loopdim - 1) reach *= K.mm.loopsize;
    sh.hi[d] = std::min<int64_t>(P.ranges[d][1], P.ranges[d][0] + reach - 1);
  }
  // inbounds=true is @inbounds over the body (src/ParallelKernel/parallel.jl:209): the caller
  // asserts that every access is inside its array and no check is emitted
  std::string why;
  v.checked = !K.opt.inbounds && !prove_in_bounds(K, sh, memopt, &why);
  // 32-bit per-thread byte offsets (Variant::off32) for plain kernels whose launch box starts at
  // index 1 or above in every dimension and whose fields' byte offsets over the box stay < 2^31.
  // They are used by x groups only: with one cell per thread the scalar bases add scalar
  // instructions to every wavefront, and in kernels that also check exact quotients (scalar
  // masks) that made the 512^3 diffusion step 25% slower (1.078 against 0.862 ms).
  if (!memopt && !v.checked && !pad && K.opt.addr32) {
    bool ok = sh.lo[0] >= 1 && sh.lo[1] >= 1 && sh.lo[2] >= 1;
    for (int i = 0; i < nargs && ok; ++i) {
      if (args[i].kind != PS_ARG_FIELD) continue;
      const ps_field& f = args[i].field;
      const int64_t top = (sh.hi[0] - 1) + (sh.hi[1] - 1) * f.pitch + (sh.hi[2] - 1) * f.plane;
      if (extent_bytes(f) >= ((int64_t)1 << 31) || top < 0 || top * (int64_t)elsize(f.dtype) >= ((int64_t)1 << 31)) ok = false;
    }
    // 24-bit multiplies for the cell offsets: every stride and every index of the box below 2^24
    // (the products then stay below 2^31 by the test above)
    bool u24 = ok && sh.hi[0] <= (1 << 24) && sh.hi[1] <= (1 << 24) && sh.hi[2] <= (1 << 24);
    for (int i = 0; i < nargs && u24; ++i)
      if (args[i].kind == PS_ARG_FIELD && (args[i].field.pitch >= (1 << 24) || args[i].field.plane >= (1 << 24))) u24 = false;
    v.u24 = u24;
    // stride classes: fields (not CellArrays) with equal strides share one cell offset
    std::string sc(nargs, '-');
    std::vector<std::pair<int64_t, int64_t>> seen;
    for (int i = 0; i < nargs; ++i) {
      if (args[i].kind != PS_ARG_FIELD || args[i].field.ncomp > 0) continue;
      const std::pair<int64_t, int64_t> st{args[i].field.pitch, args[i].field.plane};
      size_t c = 0;
      while (c < seen.size() && seen[c] != st) ++c;
      if (c == seen.size()) seen.push_back(st);
      sc[i] = (char)('a' + (c % 26));
    }
    if (seen.size() <= 26) v.sclass = sc;
    // x groups (Variant::cx): with automatic geometry and a body of array assignments under
    // x-interval masks, each thread takes cx adjacent x cells (by default 4 when the narrowest
    // array has 2- or 4-byte elements: 8 or 16 bytes per thread, whose accesses merge into wide
    // loads; FP64 arrays keep one cell, where groups measured slower)
    bool cell = false;
    int64_t es = 8;
    for (int i = 0; i < nargs; ++i)
      if (args[i].kind == PS_ARG_FIELD) {
        cell = cell || args[i].field.ncomp > 0;
        es = std::min<int64_t>(es, elsize(args[i].field.dtype));
      }
    if (ok && !user_geom && !cell && groupable_body(K)) {
      // Cells per thread. FP32 and BFloat16: 4 x cells. FP64 (3-D groups, rows of at least 256
      // cells): z cells, twice the stencil's z reach (a thread then reads 1.5 planes per cell) and
      // at most 8, so that a z plane that several of the thread's cells read is one load; and 2 x
      // cells with them when the stencil reaches 2 cells along x. Measured at the sustained clock:
      // diffusion3D 512^3 (reaches 2 and 2) 0.860 ms per step with one cell, 0.739 / 0.684 /
      // 0.704 with 2 / 4 / 8 z cells, 0.652 with 2 x 4; acoustic3D 384^3 (reaches 1 and 1)
      // 1.026, 0.961 with 2 z cells, 0.998 with 4, 1.001 with 2 x 2. FP32 and BFloat16 groups
      // were 10-15% slower with z cells, and Stokes3D at 127^3 (rows of 127 cells: the edge block
      // column and the threads before the first interior group double the waves) 4% slower.
      const bool fp64 = es >= 8 && len[0] >= 256 && groupable_body(K, true);
      int cz = 1;
      if (K.opt.cells_z != 0) cz = K.opt.cells_z;
      else if (fp64) {
        const int reach = idx_span(K, 2);
        cz = reach == 0 ? 1 : reach == 1 ? 2 : reach == 2 ? 4 : 8;
      }
      int cx = K.opt.cells_x;
      if (cx == 0) cx = es < 8 ? 4 : fp64 && cz > 1 && idx_span(K, 0) >= 2 ? 2 : 1;
      if (len[0] < 2 * (int64_t)cx) cx = 1;
      if (cz > 1 && (P.nthreads[2] != 1 || len[2] < 2 * (int64_t)cz || !groupable_body(K, true))) cz = 1;
      if (cx > 1 || cz > 1) {
        v.cx = cx;
        v.cz = cz;
        v.off32 = true;
        // Block shape of the launch: the reference's thread count per block, bx threads along x,
        // by default 128 cells along x per block row for 3-D launches (4 cells per thread: 32
        // threads, two y rows per wavefront) and 256 for 2-D ones (64 threads). Measured on
        // acoustic3D 384^3 and acoustic2D 4096^2 over 1-8 cells per thread and 16-64 threads at
        // the sustained clock: the fastest in FP32 and BFloat16.
        const int64_t ng = (len[0] + cx - 1) / cx;
        const int32_t txy = P.nthreads[0] * P.nthreads[1];
        auto pow2_ceil = [](int64_t n) { int64_t p = 1; while (p < n) p <<= 1; return p; };
        // (FP64 groups with z cells: 32 threads of 2 x cells, measured 0.652 against 0.675 ms
        // with 64 on the diffusion3D step at 512^3)
        int64_t bx = K.opt.nthreads_x_groups ? K.opt.nthreads_x_groups
                     : cx > 1 ? (cz > 1 ? 32 : std::max<int64_t>(1, (len[2] > 1 ? 128 : 256) / cx))
                              : (int64_t)P.nthreads[0];
        bx = std::min<int64_t>({bx, (int64_t)txy, pow2_ceil(ng)});
        int64_t by = std::max<int64_t>(1, txy / bx);
        if (by > pow2_ceil(len[1])) {
          by = pow2_ceil(len[1]);
          bx = std::max<int64_t>(1, txy / by);
        }
        if ((len[1] + by - 1) / by > 65535) {
          bx = P.nthreads[0];
          by = P.nthreads[1];
        }
        P.lthreads[0] = (int32_t)bx;
        P.lthreads[1] = (int32_t)by;
        // room for the threads before the first interior group (the groups start on a 128-byte
        // line: up to 128 / es / cx of them) and one block column more: its threads take the cells
        // at the row ends, outside the interior groups (the kernel computes both)
        const int64_t lead = std::max<int64_t>(1, 128 / (es * cx));
        P.lblocks[0] = (int32_t)((ng + lead + bx - 1) / bx + 1);
        P.lblocks[1] = (int32_t)((len[1] + by - 1) / by);
        if (cz > 1) P.lblocks[2] = (int32_t)((len[2] + cz - 1) / cz);  // blocks of one z slab (nthreads[2] == 1)
      }
    }
  }
  // XCD spread of x-group launches: workgroups go to the 8 XCDs round-robin in linear block
  // order (x fastest), each XCD with its own L2. With an even number of block columns along x an
  // XCD only ever gets some of the columns (all of one column when there are 8), so that its L2
  // and its share of the memory requests see a few narrow x ranges of every row; an odd number
  // rotates the columns over the XCDs from one block row to the next. An even x grid therefore
  // gets one more block column, whose threads find no cells and return. Measured on acoustic3D
  // 384^3 with x groups: 10-35% faster where the grid was even. One-cell-per-thread launches keep
  // the reference's grid (diffusion3D 512^3 FP64, 8 columns, ran 3% slower with 9), and so do
  // memopt launches (diffusion3D 512^3: 2 columns, 3% slower with 3, and 2-5% slower with the
  // blocks of each XCD taken from one contiguous eighth of the grid, xcd_remap).
  if ((v.cx > 1 || v.cz > 1) && K.opt.xcd_spread && P.lblocks[2] > 1 && P.lblocks[0] % 2 == 0) P.lblocks[0] += 1;
  P.v = v;
  return P;
}

struct LaunchEntry {
  Plan P;
  std::shared_ptr<Module> m;  // null for an empty launch box
  bool pad = false;           // padding=true: fields are passed as their parents and view boxes
};

// Everything plan_launch and the variant depend on: argument kinds and types, field shapes and
// pitches, integer scalar values (bounds proof), the division mode of float scalars, which
// field arguments overlap in memory (aliasing), and the launch request.
std::string launch_key(const ps_arg* args, int nargs, const ps_launch* L) {
  std::string k;
  k.reserve(32 + 48 * (size_t)nargs);
  auto put = [&](const void* p, size_t n) { k.append((const char*)p, n); };
  for (int i = 0; i < nargs; ++i) {
    const ps_arg& a = args[i];
    put(&a.kind, 4);
    put(&a.dtype, 4);
    if (a.kind == PS_ARG_FIELD) {
      put(a.field.dims, 24);
      put(&a.field.pitch, 8);
      put(&a.field.plane, 8);
      put(&a.field.ndims, 4);
      if (a.field.pdims[0]) {  // a view: its place in the parent (padding=true kernels address the parent)
        put(a.field.voff, 24);
        put(a.field.pdims, 24);
      }
      if (a.field.ncomp > 0) {
        put(&a.field.ncomp, 4);
        put(a.field.celldims, 16);
        put(&a.field.blocklen, 8);
        if (a.field.cellnames) k.append(a.field.cellnames);
        k.push_back(';');
      }
    } else if (a.dtype == PS_F64 || a.dtype == PS_F32) {
      char m = 'x';
      if (a.dtype == PS_F64) m = rcp_safe64(a.f) ? (a.f > 0 ? 'p' : 'n') : 'x';
      else m = rcp_safe32((float)a.f) ? ((float)a.f > 0 ? 'p' : 'n') : 'x';
      k.push_back(m);
    } else {
      put(&a.i, 8);
    }
  }
  for (int i = 0; i < nargs; ++i) {
    if (args[i].kind != PS_ARG_FIELD) continue;
    const ps_field& a = args[i].field;
    const char* a0 = (const char*)a.ptr;
    const char* a1 = a0 + extent_bytes(a);
    for (int j = i + 1; j < nargs; ++j) {
      if (args[j].kind != PS_ARG_FIELD) continue;
      const ps_field& b = args[j].field;
      const char* b0 = (const char*)b.ptr;
      const char* b1 = b0 + extent_bytes(b);
      if (a0 < b1 && b0 < a1) {
        int16_t pr[2] = {(int16_t)i, (int16_t)j};
        put(pr, 4);
      }
    }
  }
  k.push_back('|');
  if (L) {
    put(L->ranges, sizeof(L->ranges));
    put(&L->has_ranges, 4);
    put(L->nthreads, 12);
    put(L->nblocks, 12);
    put(&L->memopt, 4);
    put(&L->shmem, 8);
  }
  return k;
}

}  // namespace

extern "C" int ps_kernel_geometry(ps_kernel k, const ps_arg* args, int nargs, const ps_launch* launch, int32_t nthreads[3], int32_t nblocks[3],
                                  int64_t ranges[3][2], int64_t* lds_bytes) {
  PS_GUARD_BEGIN
  KernelObj* ko = (KernelObj*)k;
  std::vector<ArgKind> kinds;
  std::string bits;
  const Kernel& K = kernel_for(ko, args, nargs, &bits);
  Plan P = plan_launch(K, args, nargs, launch, kinds, bits);
  for (int d = 0; d < 3; ++d) {
    nthreads[d] = P.nthreads[d];
    nblocks[d] = P.nblocks[d];
    ranges[d][0] = P.ranges[d][0];
    ranges[d][1] = P.ranges[d][1];
  }
  if (lds_bytes) *lds_bytes = 0;
  return (P.memopt ? 1 : 0) | (P.v.checked ? 2 : 0) | (P.v.idx64 ? 4 : 0);
  PS_GUARD_END
}

extern "C" int ps_kernel_precompile_launch(ps_kernel k, const ps_arg* args, int nargs, const ps_launch* launch) {
  PS_GUARD_BEGIN
  KernelObj* ko = (KernelObj*)k;
  std::vector<ArgKind> kinds;
  std::string bits;
  const Kernel& K = kernel_for(ko, args, nargs, &bits);
  Plan P = plan_launch(K, args, nargs, launch, kinds, bits);
  get_module(ko, K, kinds, P.v, false);
  return 0;
  PS_GUARD_END
}

extern "C" int ps_kernel_launch(ps_kernel k, const ps_arg* args, int nargs, const ps_launch* launch) {
  if (int r = require_init()) return r;
  PS_GUARD_BEGIN
  KernelObj* ko = (KernelObj*)k;
  for (int i = 0; i < nargs; ++i) {
    const ps_field& f = args[i].field;
    if (args[i].kind == PS_ARG_FIELD && !f.ptr && f.dims[0] * f.dims[1] * f.dims[2] > 0)
      return fail(PS_EARGUMENT, "argument " + std::to_string(i + 1) + " of kernel " + ko->K.def.name +
                                    ": the field has no device memory (it was freed or never allocated)");
  }
  std::string key = launch_key(args, nargs, launch);
  std::shared_ptr<LaunchEntry> le;
  {
    std::lock_guard<std::mutex> g(ko->lmu);
    auto it = ko->launches.find(key);
    if (it != ko->launches.end()) le = it->second;
  }
  if (!le) {
    le = std::make_shared<LaunchEntry>();
    std::vector<ArgKind> kinds;
    std::string bits;
    const Kernel& K = kernel_for(ko, args, nargs, &bits);
    le->P = plan_launch(K, args, nargs, launch, kinds, bits);
    le->pad = K.opt.padding;
    bool empty = false;
    for (int d = 0; d < 3; ++d)
      if (le->P.nblocks[d] == 0) empty = true;
    if (!empty) le->m = get_module(ko, K, kinds, le->P.v, true);
    std::lock_guard<std::mutex> g(ko->lmu);
    if (ko->launches.size() >= 256) ko->launches.clear();
    ko->launches[key] = le;
  }
  if (!le->m) return 0;  // empty range
  const Plan& P = le->P;
  const auto& m = le->m;
  // pack parameters
  std::vector<unsigned long long> slots(m->em.params.size());
  std::vector<void*> ptrs(m->em.params.size());
  for (size_t i = 0; i < m->em.params.size(); ++i) {
    const auto& p = m->em.params[i];
    unsigned long long v = 0;
    long long iv = 0;
    bool isint = true;
    switch (p.kind) {
      case 0: {
        void* fp = le->pad ? parent_ptr(args[p.arg].field) : args[p.arg].field.ptr;
        std::memcpy(&v, &fp, 8);
        isint = false;
        break;
      }
      case 1: {
        const ps_field& f = args[p.arg].field;
        if (p.which < 3) iv = le->pad ? pdim(f, p.which) : f.dims[p.which];
        else iv = p.which == 3 ? f.pitch : f.plane;
        break;
      }
      case 10: {  // view box in the parent: first (which 0..2) and last (3..5) index
        const ps_field& f = args[p.arg].field;
        const int d = p.which % 3;
        iv = p.which < 3 ? voffset(f, d) + 1 : voffset(f, d) + f.dims[d];
        break;
      }
      case 2: {
        const ps_arg& a = args[p.arg];
        if (a.dtype == PS_F64) {
          std::memcpy(&v, &a.f, 8);
          isint = false;
        } else if (a.dtype == PS_F32) {
          float fv = (float)a.f;
          std::memcpy(&v, &fv, 4);
          isint = false;
        } else if (a.dtype == PS_BF16) {
          throw PsError(PS_EARGUMENT, "BFloat16 scalars are not supported; pass Float32");
        } else {
          iv = a.dtype == PS_BOOL ? (a.i != 0) : a.i;
        }
        break;
      }
      case 3: iv = P.ranges[p.which][0]; break;
      case 4: iv = P.ranges[p.which][1] - P.ranges[p.which][0] + 1; break;
      case 5: {
        void* e = (void*)S().d_err;
        std::memcpy(&v, &e, 8);
        isint = false;
        break;
      }
      case 6: {  // RN(1/b) of a uniform divisor, 0 when the fast division is not exact for b
        const ps_arg& a = args[p.arg];
        if (p.dtype == PS_F64) {
          double b = std::fabs(a.f), r = rcp_safe64(b) ? 1.0 / b : 0.0;
          std::memcpy(&v, &r, 8);
        } else {
          float b = std::fabs((float)a.f), r = rcp_safe32(b) ? 1.0f / b : 0.0f;
          std::memcpy(&v, &r, 4);
        }
        isint = false;
        break;
      }
      case 11: {  // host-evaluated scalar subexpression
        const auto& up = m->em.uprogs[p.arg];
        if (up.ty == 2) {
          double d = eval_uprog_f64(up, args);
          std::memcpy(&v, &d, 8);
        } else {
          float f = eval_uprog_f32(up, args);
          std::memcpy(&v, &f, 4);
        }
        isint = false;
        break;
      }
      case 8: iv = launch ? launch->shmem : 0; break;
      case 9: iv = args[p.arg].field.blocklen; break;
      case 7: {  // fastmath: RN(1/b) of a uniform divisor
        const ps_arg& a = args[p.arg];
        if (p.dtype == PS_F64) {
          double r = 1.0 / a.f;
          std::memcpy(&v, &r, 8);
        } else {
          float r = 1.0f / (float)a.f;
          std::memcpy(&v, &r, 4);
        }
        isint = false;
        break;
      }
    }
    if (isint) {
      if (p.bytes == 4) {
        int32_t x = (int32_t)iv;
        std::memcpy(&v, &x, 4);
      } else {
        std::memcpy(&v, &iv, 8);
      }
    }
    slots[i] = v;
    ptrs[i] = &slots[i];
  }
  hipStream_t st = stream_of(launch ? launch->stream : nullptr);
  const unsigned shmem = launch ? (unsigned)launch->shmem : 0u;
  if (shmem) {
    static int lds_max = -1;
    if (lds_max < 0) PS_HIP(hipDeviceGetAttribute(&lds_max, hipDeviceAttributeMaxSharedMemoryPerBlock, S().device));
    if ((int64_t)shmem + m->em.lds_bytes > lds_max)
      return fail(PS_EARGUMENT, "@parallel <kernelcall>: shmem=" + std::to_string(shmem) + " bytes exceeds the " + std::to_string(lds_max) +
                                    " bytes of LDS a workgroup can use");
  }
  if (P.v.checked) g_err_armed.store(true, std::memory_order_relaxed);
  hipError_t e = hipModuleLaunchKernel(m->fn, (unsigned)P.lblocks[0], (unsigned)P.lblocks[1], (unsigned)P.lblocks[2], (unsigned)P.lthreads[0],
                                       (unsigned)P.lthreads[1], (unsigned)P.lthreads[2], shmem, st, ptrs.data(), nullptr);
  if (e != hipSuccess) return fail(PS_EHIP, std::string("hipModuleLaunchKernel: ") + hipGetErrorString(e));
  if (!(launch && launch->async)) {
    PS_HIP(hipStreamSynchronize(st));
    if (P.v.checked) return check_err_flag(st);
  }
  return 0;
  PS_GUARD_END
}
