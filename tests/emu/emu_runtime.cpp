/*
 * Host side of the CPU emulation of HIP (test infrastructure): device
 * allocations as separate heap blocks (so AddressSanitizer bounds every
 * buffer), a registry of them for pointer-argument checks, synchronous
 * copies, and the block executor: one block at a time, its threads run
 * concurrently on a persistent pool with a real barrier for __syncthreads.
 */
#include <hip/hip_runtime.h>

#include <sys/mman.h>

#include <barrier>
#include <condition_variable>
#include <map>
#include <memory>
#include <mutex>
#include <set>
#include <string>
#include <thread>
#include <vector>

thread_local dim3 threadIdx;
thread_local dim3 blockIdx;
dim3 blockDim;
dim3 gridDim;

namespace {

std::mutex g_reg_mu;
std::map<uintptr_t, size_t> g_reg;   /* base -> size */

void reg_add(void *p, size_t n)
{
  std::lock_guard<std::mutex> l(g_reg_mu);
  g_reg[(uintptr_t) p] = n;
}

bool reg_del(void *p)
{
  std::lock_guard<std::mutex> l(g_reg_mu);
  return g_reg.erase((uintptr_t) p) == 1;
}

const char *g_kernel = "";

/* per-block state shared by the pool threads */
struct BlockState {
  std::unique_ptr<std::barrier<>> bar;
  std::vector<double> xd;
  std::vector<float> xf;
  std::vector<int> xi;
  int nthreads = 0;
};
BlockState g_blk;

int flat_tid(void)
{
  return (int) (threadIdx.x + threadIdx.y * blockDim.x
    + threadIdx.z * blockDim.x * blockDim.y);
}

class Pool {
 public:
  void run(int k, const std::function<void(int)> &job)
  {
    ensure(k);
    {
      std::unique_lock<std::mutex> l(mu_);
      job_ = &job;
      active_ = k;
      remaining_ = k;
      gen_++;
    }
    cv_.notify_all();
    std::unique_lock<std::mutex> l(mu_);
    done_.wait(l, [&] { return remaining_ == 0; });
  }

 private:
  void ensure(int k)
  {
    while((int) th_.size() < k) {
      int id = (int) th_.size();
      th_.emplace_back([this, id] { worker(id); });
      th_.back().detach();
    }
  }
  void worker(int id)
  {
    uint64_t seen = 0;
    for(;;) {
      std::unique_lock<std::mutex> l(mu_);
      cv_.wait(l, [&] { return gen_ != seen; });
      seen = gen_;
      if(id >= active_) continue;
      const std::function<void(int)> *j = job_;
      l.unlock();
      (*j)(id);
      l.lock();
      if(--remaining_ == 0) done_.notify_all();
    }
  }
  std::vector<std::thread> th_;
  std::mutex mu_;
  std::condition_variable cv_, done_;
  uint64_t gen_ = 0;
  int active_ = 0, remaining_ = 0;
  const std::function<void(int)> *job_ = nullptr;
};

Pool *g_pool = nullptr;

template<typename T>
T shfl(std::vector<T> &buf, T v, unsigned delta, int width)
{
  int tid = flat_tid();
  buf[tid] = v;
  g_blk.bar->arrive_and_wait();
  int lane = tid % width;
  int base = tid - lane;
  T out = v;
  if(lane + (int) delta < width && base + lane + (int) delta < g_blk.nthreads)
    out = buf[base + lane + delta];
  g_blk.bar->arrive_and_wait();
  return out;
}

template<typename T>
T shfl_idx(std::vector<T> &buf, T v, int src, int width)
{
  int tid = flat_tid();
  buf[tid] = v;
  g_blk.bar->arrive_and_wait();
  int lane = tid % width;
  int base = tid - lane;
  int s = ((src % width) + width) % width;
  T out = v;
  if(base + s < g_blk.nthreads) out = buf[base + s];
  g_blk.bar->arrive_and_wait();
  (void) lane;
  return out;
}

}  // namespace

namespace bb_emu {

void syncthreads(void) { g_blk.bar->arrive_and_wait(); }

double shfl_down_d(double v, unsigned d, int w) { return shfl(g_blk.xd, v, d, w); }
float shfl_down_f(float v, unsigned d, int w) { return shfl(g_blk.xf, v, d, w); }
int shfl_down_i(int v, unsigned d, int w) { return shfl(g_blk.xi, v, d, w); }
double shfl_idx_d(double v, int src, int w) { return shfl_idx(g_blk.xd, v, src, w); }
int shfl_idx_i(int v, int src, int w) { return shfl_idx(g_blk.xi, v, src, w); }

unsigned long long ballot(int p)
{
  int tid = flat_tid();
  g_blk.xi[tid] = p != 0;
  g_blk.bar->arrive_and_wait();
  int base = tid - tid % 64;
  unsigned long long m = 0;
  for(int l = 0; l < 64 && base + l < g_blk.nthreads; l++)
    if(g_blk.xi[base + l]) m |= 1ull << l;
  g_blk.bar->arrive_and_wait();
  return m;
}

bool is_device_range(const void *p, size_t bytes)
{
  if(p == nullptr) return true;
  std::lock_guard<std::mutex> l(g_reg_mu);
  uintptr_t a = (uintptr_t) p;
  auto it = g_reg.upper_bound(a);
  if(it == g_reg.begin()) return false;
  --it;
  return a >= it->first && a + bytes <= it->first + (it->second ? it->second : 1);
}

void report_bad_pointer(const char *kernel, int arg, const void *p)
{
  fprintf(stderr, "bb_emu: kernel %s argument %d = %p is not a device "
    "allocation\n", kernel, arg, p);
  fflush(stderr);
  abort();
}

void run_grid(const char *name, dim3 g, dim3 b, const std::function<void()> &body)
{
  if(!g_pool) g_pool = new Pool();
  long nth = (long) b.x * b.y * b.z;
  if(nth <= 0 || nth > 1024 || g.x == 0 || g.y == 0 || g.z == 0
     || g.y > 65535 || g.z > 65535) {
    fprintf(stderr, "bb_emu: kernel %s has an invalid launch shape "
      "grid (%u,%u,%u) block (%u,%u,%u)\n", name, g.x, g.y, g.z, b.x, b.y, b.z);
    fflush(stderr);
    abort();
  }
  g_kernel = name;
  gridDim = g;
  blockDim = b;
  g_blk.nthreads = (int) nth;
  g_blk.xd.assign(nth, 0.);
  g_blk.xf.assign(nth, 0.f);
  g_blk.xi.assign(nth, 0);
  for(unsigned bz = 0; bz < g.z; bz++)
  for(unsigned by = 0; by < g.y; by++)
  for(unsigned bx = 0; bx < g.x; bx++) {
    g_blk.bar.reset(new std::barrier<>(nth));
    dim3 bi(bx, by, bz);
    std::function<void(int)> job = [&](int t) {
      blockIdx = bi;
      threadIdx = dim3(t % b.x, (t / b.x) % b.y, t / (b.x * b.y));
      body();
      g_blk.bar->arrive_and_drop();
    };
    g_pool->run((int) nth, job);
  }
}

}  // namespace bb_emu

/* ---- runtime API ---- */
hipError_t hipGetDeviceCount(int *n) { *n = 1; return hipSuccess; }
hipError_t hipSetDevice(int) { return hipSuccess; }
hipError_t hipGetDevice(int *d) { *d = 0; return hipSuccess; }
hipError_t hipGetDeviceProperties(hipDeviceProp_t *p, int)
{
  memset(p, 0, sizeof(*p));
  snprintf(p->name, sizeof(p->name), "CPU emulation (test build)");
  snprintf(p->gcnArchName, sizeof(p->gcnArchName), "emu");
  p->multiProcessorCount = 1;
  return hipSuccess;
}
hipError_t hipStreamCreateWithFlags(hipStream_t *s, unsigned)
{
  *s = (hipStream_t) 0x1;
  return hipSuccess;
}
hipError_t hipStreamSynchronize(hipStream_t) { return hipSuccess; }
hipError_t hipDeviceSynchronize(void) { return hipSuccess; }

/* BB_EMU_GUARD=right (or left) places every device allocation against a
 * 1 GiB band of inaccessible address space on both sides, its end (or its
 * start) flush with a page boundary, so that any access past it -- however
 * far -- faults at once instead of landing in another heap block. */
static int guard_mode(void)
{
  static int m = -1;
  if(m < 0) {
    const char *e = getenv("BB_EMU_GUARD");
    m = !e ? 0 : strcmp(e, "right") == 0 ? 1 : strcmp(e, "left") == 0 ? 2 : 0;
  }
  return m;
}

static const size_t GUARD = (size_t) 1 << 30;
static std::map<uintptr_t, std::pair<uintptr_t, size_t>> g_maps;  /* ptr -> (map, len) */

static void *guarded_alloc(size_t n)
{
  size_t pg = 4096;
  size_t body = (n + pg - 1) / pg * pg;
  if(body == 0) body = pg;
  size_t len = body + 2 * GUARD;
  void *m = mmap(NULL, len, PROT_NONE, MAP_PRIVATE | MAP_ANONYMOUS | MAP_NORESERVE,
    -1, 0);
  if(m == MAP_FAILED) return NULL;
  char *b = (char *) m + GUARD;
  if(mprotect(b, body, PROT_READ | PROT_WRITE) != 0) return NULL;
  char *ptr = b;
  if(guard_mode() == 1) {
    size_t r = (n + 15) / 16 * 16;
    ptr = b + body - r;
  }
  std::lock_guard<std::mutex> l(g_reg_mu);
  g_maps[(uintptr_t) ptr] = std::make_pair((uintptr_t) m, len);
  return ptr;
}

static bool guarded_free(void *p)
{
  std::pair<uintptr_t, size_t> mp;
  {
    std::lock_guard<std::mutex> l(g_reg_mu);
    auto it = g_maps.find((uintptr_t) p);
    if(it == g_maps.end()) return false;
    mp = it->second;
    g_maps.erase(it);
  }
  munmap((void *) mp.first, mp.second);
  return true;
}

hipError_t hipMalloc(void **p, size_t n)
{
  *p = guard_mode() ? guarded_alloc(n) : malloc(n ? n : 1);
  if(!*p) return hipErrorOutOfMemory;
  /* fill never-written device memory with a pattern: 0xff bytes by
   * default (NaN doubles, -1 ints); BB_EMU_FILL=zero or =random tries the
   * other contents a fresh GPU allocation can have */
  static int mode = -1;
  if(mode < 0) {
    const char *e = getenv("BB_EMU_FILL");
    mode = !e ? 0 : strcmp(e, "zero") == 0 ? 1 : strcmp(e, "random") == 0 ? 2 : 0;
  }
  if(mode == 0) memset(*p, 0xff, n);
  else if(mode == 1) memset(*p, 0, n);
  else {
    static uint64_t st = 0x9E3779B97F4A7C15ull;
    unsigned char *b = (unsigned char *) *p;
    for(size_t q = 0; q < n; q++) {
      st ^= st << 13; st ^= st >> 7; st ^= st << 17;
      b[q] = (unsigned char) st;
    }
  }
  reg_add(*p, n);
  return hipSuccess;
}
hipError_t hipFree(void *p)
{
  if(!p) return hipSuccess;
  if(!reg_del(p)) {
    fprintf(stderr, "bb_emu: hipFree of %p, not a device allocation\n", p);
    abort();
  }
  if(!guarded_free(p)) free(p);
  return hipSuccess;
}
static std::set<uintptr_t> g_host_allocs;

hipError_t bb_emu_symbol_address(void **p, const void *sym, size_t bytes)
{
  static std::set<uintptr_t> seen;
  {
    std::lock_guard<std::mutex> l(g_reg_mu);
    if(seen.count((uintptr_t) sym)) { *p = (void *) sym; return hipSuccess; }
    seen.insert((uintptr_t) sym);
  }
  reg_add((void *) sym, bytes);
  *p = (void *) sym;
  return hipSuccess;
}

hipError_t hipHostMalloc(void **p, size_t n, unsigned)
{
  *p = malloc(n ? n : 1);
  reg_add(*p, n);
  {
    std::lock_guard<std::mutex> l(g_reg_mu);
    g_host_allocs.insert((uintptr_t) *p);
  }
  return *p ? hipSuccess : hipErrorOutOfMemory;
}
hipError_t hipHostFree(void *p)
{
  reg_del(p);
  {
    std::lock_guard<std::mutex> l(g_reg_mu);
    g_host_allocs.erase((uintptr_t) p);
  }
  free(p);
  return hipSuccess;
}

/* ---- launch capture for the machine-code replay (tests/isa) ---- */
namespace {
std::map<std::string, int> g_cap_count;
int g_cap_seq = 0;
std::string g_cap_dir;

const char *cap_dir(void)
{
  static int init = 0;
  if(!init) {
    init = 1;
    const char *e = getenv("BB_EMU_ISA_DUMP");
    if(e) g_cap_dir = e;
  }
  return g_cap_dir.empty() ? nullptr : g_cap_dir.c_str();
}

/* device allocations only: pinned host buffers are never kernel operands */
std::vector<std::pair<uintptr_t, size_t>> cap_allocs(void)
{
  std::vector<std::pair<uintptr_t, size_t>> out;
  std::lock_guard<std::mutex> l(g_reg_mu);
  for(auto &kv : g_reg)
    if(!g_host_allocs.count(kv.first)) out.push_back(kv);
  return out;
}
}

namespace bb_emu {
bool isa_capture_wanted(const char *name)
{
  if(!cap_dir()) return false;
  const char *mx = getenv("BB_EMU_ISA_MAX");
  int lim = mx ? atoi(mx) : 2;
  const char *only = getenv("BB_EMU_ISA_KERNELS");
  if(only && *only) {
    std::string list = std::string(",") + only + ",";
    if(list.find(std::string(",") + name + ",") == std::string::npos) return false;
  }
  return g_cap_count[name] < lim;
}

void isa_capture(const char *name, const char *ftype, dim3 g, dim3 b,
  const std::vector<unsigned char> &args, int phase)
{
  static std::string stem;
  static std::vector<std::pair<uintptr_t, size_t>> al;
  if(phase == 0) {
    char buf[512];
    std::string nm = name;
    for(auto &c : nm) if(!isalnum((unsigned char) c) && c != '_') c = '_';
    snprintf(buf, sizeof(buf), "%s/%05d_%s", cap_dir(), g_cap_seq++, nm.c_str());
    stem = buf;
    al = cap_allocs();
    FILE *f = fopen((stem + ".json").c_str(), "w");
    fprintf(f, "{\"kernel\": \"%s\", \"ftype\": \"%s\", \"grid\": [%u, %u, %u], "
      "\"block\": [%u, %u, %u], \"args\": \"", name, ftype, g.x, g.y, g.z, b.x, b.y, b.z);
    for(unsigned char c : args) fprintf(f, "%02x", c);
    fprintf(f, "\", \"allocs\": [");
    for(size_t i = 0; i < al.size(); i++)
      fprintf(f, "%s[%llu, %zu]", i ? ", " : "", (unsigned long long) al[i].first,
        al[i].second);
    fprintf(f, "]}\n");
    fclose(f);
  }
  FILE *f = fopen((stem + (phase == 0 ? ".before" : ".after")).c_str(), "wb");
  for(auto &a : al) fwrite((const void *) a.first, 1, a.second, f);
  fclose(f);
  if(phase == 1) g_cap_count[name]++;
}
}

static void check_copy(void *d, const void *s, size_t n, hipMemcpyKind k)
{
  bool dd = k == hipMemcpyHostToDevice || k == hipMemcpyDeviceToDevice;
  bool sd = k == hipMemcpyDeviceToHost || k == hipMemcpyDeviceToDevice;
  if((dd && !bb_emu::is_device_range(d, n)) || (sd && !bb_emu::is_device_range(s, n))) {
    fprintf(stderr, "bb_emu: copy of %zu bytes kind %d between %p and %p is "
      "out of a device allocation\n", n, (int) k, s, d);
    fflush(stderr);
    abort();
  }
}

hipError_t hipMemcpy(void *d, const void *s, size_t n, hipMemcpyKind k)
{
  check_copy(d, s, n, k);
  memmove(d, s, n);
  return hipSuccess;
}
hipError_t hipMemcpyAsync(void *d, const void *s, size_t n, hipMemcpyKind k,
  hipStream_t)
{
  return hipMemcpy(d, s, n, k);
}
hipError_t hipMemsetAsync(void *d, int v, size_t n, hipStream_t)
{
  if(!bb_emu::is_device_range(d, n)) {
    fprintf(stderr, "bb_emu: memset of %zu bytes at %p is out of a device "
      "allocation\n", n, d);
    abort();
  }
  memset(d, v, n);
  return hipSuccess;
}
struct bb_emu_event { int dummy; };
hipError_t hipEventCreate(hipEvent_t *e) { *e = new bb_emu_event(); return hipSuccess; }
hipError_t hipEventRecord(hipEvent_t, hipStream_t) { return hipSuccess; }
hipError_t hipEventSynchronize(hipEvent_t) { return hipSuccess; }
hipError_t hipEventDestroy(hipEvent_t e) { delete e; return hipSuccess; }
hipError_t hipEventElapsedTime(float *ms, hipEvent_t, hipEvent_t) { *ms = 0.f; return hipSuccess; }
hipError_t hipGetLastError(void) { return hipSuccess; }
const char *hipGetErrorString(hipError_t) { return "emulated error"; }
