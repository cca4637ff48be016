// ConvectionField.hpp -- the field type of the drop-in C++ host API.
//
// Mirrors the part of the reference's ConvectionField (src/Convection/Convection.h:13-20;
// the STELLA DataFieldCUDA API) that src/Parareal/main.cpp and Parareal.h use:
//   Init(name, domain, kboundary), calculationDomain(), boundary(), storage().pStorageBase(),
//   storage().paddedSize(), operator()(i, j, k) on the host copy, SynchronizeHostStorage(),
//   SynchronizeDeviceStorage(), SwapWith(), copy.
// Storage is the reference GPU layout (KJI order, i padded to 32 doubles, 2-cell halo), on
// the device through the C ABI; pStorageBase() is the device address, as on the reference's
// CUDA backend (DataFieldCUDAStorage.h:123).  A host mirror holds the same bytes for element
// access and follows the device lazily: element writes mark the mirror newer, and the next
// device use uploads it; handing out the device address marks the device newer, and the
// next element access downloads it.  So call sites that skip the explicit Synchronize*
// calls (the reference's CPU-backend form of main.cpp) still see consistent data.
#pragma once

#include <cstdlib>
#include <cstring>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "../stencilparareal.h"

namespace sp {

inline void check(int rc)
{
    if (rc != SP_OK) throw std::runtime_error(std::string("stencilparareal: ") + sp_last_error());
}

// One library context per process (device 0 unless SP_DEVICE is set before first use).
inline sp_ctx* context()
{
    static sp_ctx* ctx = [] {
        sp_ctx* c = nullptr;
        int dev = 0;
        if (const char* d = std::getenv("SP_DEVICE")) dev = std::atoi(d);
        check(sp_ctx_create(dev, &c));
        return c;
    }();
    return ctx;
}

// Halo width of every ConvectionField (src/Convection/Convection.h:13).
static const int convectionBoundaryLines = 2;

struct IJKSize {
    int i = 0, j = 0, k = 0;
    void Init(int isize, int jsize, int ksize)
    {
        i = isize;
        j = jsize;
        k = ksize;
    }
    int iSize() const { return i; }
    int jSize() const { return j; }
    int kSize() const { return k; }
};

struct KBoundary {
    int m = -convectionBoundaryLines, p = convectionBoundaryLines;
    void Init(int kminus, int kplus)
    {
        m = kminus;
        p = kplus;
    }
    int kMinusOffset() const { return m; }
    int kPlusOffset() const { return p; }
};

// Halo offsets of a field on each side (negative on the minus sides), as STELLA reports them.
struct IJKBoundary {
    int im = 0, ip = 0, jm = 0, jp = 0, km = 0, kp = 0;
    void Init(int iMinus, int iPlus, int jMinus, int jPlus, int kMinus, int kPlus)
    {
        im = iMinus;
        ip = iPlus;
        jm = jMinus;
        jp = jPlus;
        km = kMinus;
        kp = kPlus;
    }
    int iMinusOffset() const { return im; }
    int iPlusOffset() const { return ip; }
    int jMinusOffset() const { return jm; }
    int jPlusOffset() const { return jp; }
    int kMinusOffset() const { return km; }
    int kPlusOffset() const { return kp; }
};

class ConvectionField {
public:
    // What field.storage() returns: the device storage's base address and padded extent.
    class Storage {
    public:
        explicit Storage(ConvectionField& f) : f_(f) {}
        // first element of the allocation (the halo corner), device memory
        double* pStorageBase() const { return static_cast<double*>(f_.device()); }
        // allocated extent: i padded to 32 doubles incl. halo, j and k incl. halo
        IJKSize paddedSize() const
        {
            IJKSize s;
            const sp_layout& l = f_.layout_;
            s.Init((int)l.sj, (int)(l.sk / l.sj), l.nk + 2 * l.halo);
            return s;
        }

    private:
        ConvectionField& f_;
    };

    ConvectionField() = default;
    ConvectionField(const ConvectionField& o) { *this = o; }
    ConvectionField& operator=(const ConvectionField& o)
    {
        if (this == &o) return *this;
        release();
        if (o.dev_) {
            Init(o.name_, o.domain_, KBoundary{});
            // a full copy of the storage (the reference's copy constructor copies data)
            o.pull();
            host_ = o.host_;
            dev_newer_ = false;
            host_newer_ = true;
            push();
        }
        return *this;
    }
    ~ConvectionField() { release(); }

    void Init(const std::string& name, const IJKSize& domain, const KBoundary& /*kboundary*/)
    {
        release();
        name_ = name;
        domain_ = domain;
        check(sp_layout_stella_cuda(domain.i, domain.j, domain.k, &layout_));
        elems_ = (size_t)sp_layout_storage_elems(&layout_);
        void* p = nullptr;
        check(sp_malloc(context(), (long long)(elems_ * sizeof(double)), &p));
        dev_ = p;
        host_.assign(elems_, 0.);
        host_newer_ = true;
        dev_newer_ = false;
        push();
    }

    const std::string& name() const { return name_; }
    const IJKSize& calculationDomain() const { return domain_; }
    IJKBoundary boundary() const
    {
        const int h = layout_.halo;
        IJKBoundary b;
        b.Init(-h, h, -h, h, -h, h);
        return b;
    }
    Storage storage() { return Storage(*this); }
    const sp_layout& layout() const { return layout_; }
    // Device address for a kernel call: uploads pending element writes; the device copy is
    // then taken as the newer one.
    void* device()
    {
        push();
        dev_newer_ = true;
        return dev_;
    }
    const void* device() const
    {
        push();
        return dev_;
    }

    // host element access (interior coordinates, halo cells at -2..-1 and n..n+1)
    double& operator()(int i, int j, int k)
    {
        pull();
        host_newer_ = true;
        return host_[index(i, j, k)];
    }
    double operator()(int i, int j, int k) const
    {
        pull();
        return host_[index(i, j, k)];
    }

    // the device copy holds every element write (STELLA's DataFieldCUDA::isDeviceUpToDate)
    bool isDeviceUpToDate() const { return !host_newer_; }

    void SynchronizeHostStorage() const
    {
        dev_newer_ = true;
        pull();
    }
    void SynchronizeDeviceStorage()
    {
        host_newer_ = true;
        push();
    }

    void SwapWith(ConvectionField& o)
    {
        std::swap(name_, o.name_);
        std::swap(domain_, o.domain_);
        std::swap(layout_, o.layout_);
        std::swap(dev_, o.dev_);
        std::swap(elems_, o.elems_);
        std::swap(host_, o.host_);
        std::swap(host_newer_, o.host_newer_);
        std::swap(dev_newer_, o.dev_newer_);
    }

private:
    size_t index(int i, int j, int k) const
    {
        return (size_t)(layout_.origin + i + (long long)j * layout_.sj + (long long)k * layout_.sk);
    }
    // device -> host mirror, after the work queued on the library stream
    void pull() const
    {
        if (!dev_newer_ || !dev_) return;
        check(sp_ctx_sync(context()));
        check(sp_memcpy_d2h(context(), host_.data(), dev_, (long long)(elems_ * sizeof(double))));
        dev_newer_ = false;
    }
    // host mirror -> device (sp_memcpy_h2d is synchronous)
    void push() const
    {
        if (!host_newer_ || !dev_) return;
        check(sp_ctx_sync(context()));
        check(sp_memcpy_h2d(context(), dev_, host_.data(), (long long)(elems_ * sizeof(double))));
        host_newer_ = false;
    }
    void release()
    {
        if (dev_) (void)sp_free(context(), dev_);
        dev_ = nullptr;
    }

    std::string name_;
    IJKSize domain_;
    sp_layout layout_{};
    void* dev_ = nullptr;
    size_t elems_ = 0;
    mutable std::vector<double> host_;
    mutable bool host_newer_ = false;  // element writes not yet on the device
    mutable bool dev_newer_ = false;   // the device copy may hold newer data than the mirror
};

}  // namespace sp
