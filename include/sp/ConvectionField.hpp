// ConvectionField.hpp -- the field type of the drop-in C++ host API.
//
// Mirrors the part of the reference's ConvectionField (src/Convection/Convection.h:13-20;
// the STELLA DataFieldCUDA API) that src/Parareal/main.cpp and Parareal.h use:
//   Init(name, domain, kboundary), calculationDomain(), operator()(i, j, k) on the host
//   copy, SynchronizeHostStorage(), SynchronizeDeviceStorage(), SwapWith(), copy.
// Storage is the reference GPU layout (KJI order, i padded to 32 doubles, 2-cell halo), on
// the device through the C ABI; a host mirror holds the same bytes for element access.
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
    int m = -2, p = 2;
    void Init(int kminus, int kplus)
    {
        m = kminus;
        p = kplus;
    }
};

class ConvectionField {
public:
    ConvectionField() = default;
    ConvectionField(const ConvectionField& o) { *this = o; }
    ConvectionField& operator=(const ConvectionField& o)
    {
        if (this == &o) return *this;
        release();
        if (o.dev_) {
            Init(o.name_, o.domain_, KBoundary{});
            check(sp_ctx_sync(context()));
            // device copy of the whole storage (the reference's copy constructor copies data)
            std::vector<double> tmp(elems_);
            check(sp_memcpy_d2h(context(), tmp.data(), o.dev_, (long long)(elems_ * sizeof(double))));
            check(sp_memcpy_h2d(context(), dev_, tmp.data(), (long long)(elems_ * sizeof(double))));
            host_ = tmp;
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
        check(sp_memcpy_h2d(context(), dev_, host_.data(), (long long)(elems_ * sizeof(double))));
    }

    const std::string& name() const { return name_; }
    const IJKSize& calculationDomain() const { return domain_; }
    const sp_layout& layout() const { return layout_; }
    void* device() { return dev_; }
    const void* device() const { return dev_; }

    // host element access (interior coordinates, halo cells at -2..-1 and n..n+1)
    double& operator()(int i, int j, int k) { return host_[index(i, j, k)]; }
    double operator()(int i, int j, int k) const { return host_[index(i, j, k)]; }

    void SynchronizeHostStorage() const
    {
        check(sp_ctx_sync(context()));
        check(sp_memcpy_d2h(context(), const_cast<double*>(host_.data()), dev_, (long long)(elems_ * sizeof(double))));
    }
    void SynchronizeDeviceStorage()
    {
        check(sp_memcpy_h2d(context(), dev_, host_.data(), (long long)(elems_ * sizeof(double))));
    }

    void SwapWith(ConvectionField& o)
    {
        std::swap(name_, o.name_);
        std::swap(domain_, o.domain_);
        std::swap(layout_, o.layout_);
        std::swap(dev_, o.dev_);
        std::swap(elems_, o.elems_);
        std::swap(host_, o.host_);
    }

private:
    size_t index(int i, int j, int k) const
    {
        return (size_t)(layout_.origin + i + (long long)j * layout_.sj + (long long)k * layout_.sk);
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
    std::vector<double> host_;
};

}  // namespace sp
