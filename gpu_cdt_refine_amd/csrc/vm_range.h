// Device arrays that grow in place.  The address range of an array's
// largest size is reserved once; physical memory is mapped onto its end as
// the array grows.  Growth copies nothing, the device address never changes
// (kernels in flight, captured graphs and cached kernel arguments stay
// valid), and memory already mapped is never moved.
#pragma once

#include <hip/hip_runtime.h>

#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

namespace gcdt {

class VmRange {
 public:
  VmRange() = default;
  VmRange(const VmRange &) = delete;
  VmRange &operator=(const VmRange &) = delete;
  ~VmRange() { release(); }

  // Does the device map memory into reserved ranges?
  static bool supported(int device) {
    int v = 0;
    if (hipDeviceGetAttribute(&v, hipDeviceAttributeVirtualMemoryManagementSupported, device) !=
        hipSuccess)
      return false;
    return v != 0;
  }

  // Reserve `bytes` of device address space (rounded up to the allocation
  // granularity).  Returns false, with nothing held, if the runtime refuses.
  bool reserve(int device, size_t bytes) {
    release();
    device_ = device;
    prop_ = hipMemAllocationProp{};
    prop_.type = hipMemAllocationTypePinned;
    prop_.location.type = hipMemLocationTypeDevice;
    prop_.location.id = device;
    if (hipMemGetAllocationGranularity(&gran_, &prop_, hipMemAllocationGranularityMinimum) !=
            hipSuccess ||
        gran_ == 0)
      return false;
    // chunks of whole 2 MiB pages: a chunk that ends inside a larger page
    // leaves the next chunk's access setting refused on this runtime
    const size_t page = size_t(2) << 20;
    gran_ = (page + gran_ - 1) / gran_ * gran_;
    const size_t size = round_up(bytes < gran_ ? gran_ : bytes);
    void *p = nullptr;
    if (hipMemAddressReserve(&p, size, gran_, nullptr, 0) != hipSuccess || !p) return false;
    base_ = static_cast<char *>(p);
    reserved_ = size;
    return true;
  }

  // Make the first `bytes` of the range usable (maps one new chunk covering
  // the difference); throws if the range is too small or the map fails.
  void map_to(size_t bytes) {
    const size_t need = round_up(bytes);
    if (need <= mapped_) return;
    if (need > reserved_) throw std::runtime_error("device array beyond its reserved range");
    const size_t size = need - mapped_;
    hipMemGenericAllocationHandle_t h{};
    check(hipMemCreate(&h, size, &prop_, 0), "hipMemCreate");
    hipError_t e = hipMemMap(base_ + mapped_, size, 0, h, 0);
    if (e != hipSuccess) {
      (void)hipMemRelease(h);
      throw std::runtime_error(std::string("hipMemMap: ") + hipGetErrorString(e));
    }
    hipMemAccessDesc desc{};
    desc.location = prop_.location;
    desc.flags = hipMemAccessFlagsProtReadWrite;
    // access is set on the whole mapped range: this runtime refuses the
    // setting for a chunk that does not start the range when the chunks'
    // sizes differ
    e = hipMemSetAccess(base_, need, &desc, 1);
    if (e != hipSuccess) {
      (void)hipMemUnmap(base_ + mapped_, size);
      (void)hipMemRelease(h);
      throw std::runtime_error(std::string("hipMemSetAccess: ") + hipGetErrorString(e) +
                               " (offset " + std::to_string(mapped_) + ", " +
                               std::to_string(size) + " bytes)");
    }
    chunks_.push_back({h, size});
    mapped_ = need;
  }

  void *ptr() const { return base_; }
  size_t mapped() const { return mapped_; }

  // Unmaps and frees everything (the caller has drained the device's work).
  void release() {
    size_t off = mapped_;
    for (auto it = chunks_.rbegin(); it != chunks_.rend(); ++it) {
      off -= it->second;
      (void)hipMemUnmap(base_ + off, it->second);
      (void)hipMemRelease(it->first);
    }
    chunks_.clear();
    if (base_) (void)hipMemAddressFree(base_, reserved_);
    base_ = nullptr;
    reserved_ = mapped_ = 0;
  }

 private:
  size_t round_up(size_t b) const { return (b + gran_ - 1) / gran_ * gran_; }
  static void check(hipError_t e, const char *what) {
    if (e != hipSuccess) throw std::runtime_error(std::string(what) + ": " + hipGetErrorString(e));
  }
  char *base_ = nullptr;
  size_t reserved_ = 0, mapped_ = 0, gran_ = 0;
  int device_ = 0;
  hipMemAllocationProp prop_{};
  std::vector<std::pair<hipMemGenericAllocationHandle_t, size_t>> chunks_;
};

}  // namespace gcdt
