// MatFile.hpp -- MATLAB level-5 MAT-file writer for 3-D double fields.
//
// Mirrors the reference's MatFile (src/MatFile.h:10-28, src/MatFile.cpp:10-76): a file of
// named 3-D double arrays (interior only, i fastest, dimensions [iSize jSize kSize]),
// optionally grouped in a cell array (startCell / endCell).  The byte layout is the
// public MAT v5 format: a 128-byte header (text, subsystem offset, version 0x0100,
// endian indicator "IM"), then one miMATRIX element per array, each holding the array
// flags (mxDOUBLE_CLASS), the dimensions, the name and the miDOUBLE data, every
// sub-element padded to 8 bytes.
#pragma once

#include <cstdint>
#include <cstring>
#include <fstream>
#include <stdexcept>
#include <string>
#include <vector>

namespace sp {

class MatFile {
public:
    explicit MatFile(const std::string& fname) : os_(fname, std::ios::binary | std::ios::trunc)
    {
        if (!os_) throw std::runtime_error("MatFile: cannot open " + fname);
        char head[128];
        std::memset(head, ' ', 116);
        const char text[] = "MATLAB 5.0 MAT-file, written by stencilparareal_amd (MI355X StencilParareal)";
        std::memcpy(head, text, sizeof(text) - 1);
        std::memset(head + 116, 0, 8);  // no subsystem data
        const uint16_t version = 0x0100, endian = ('M' << 8) | 'I';
        std::memcpy(head + 124, &version, 2);
        std::memcpy(head + 126, &endian, 2);
        os_.write(head, 128);
    }
    ~MatFile() { close(); }

    // A cell array `name` of n elements; the arrays added until endCell() are its elements.
    void startCell(const std::string& name, int n)
    {
        if (cell_ >= 0) throw std::runtime_error("MatFile: nested cells are not supported");
        std::vector<char> b;
        put_u32(b, kMiMatrix);
        put_u32(b, 0);  // size, patched by endCell
        put_flags(b, kMxCell);
        const int32_t dims[2] = {n, 1};
        put_dims(b, dims, 2);
        put_name(b, name);
        cell_ = (long long)os_.tellp();
        os_.write(b.data(), (std::streamsize)b.size());
        cell_n_ = n;
        cell_added_ = 0;
    }
    void endCell()
    {
        if (cell_ < 0) return;
        if (cell_added_ != cell_n_)
            throw std::runtime_error("MatFile: cell declared " + std::to_string(cell_n_) + " elements, got " +
                                     std::to_string(cell_added_));
        const long long end = (long long)os_.tellp();
        const uint32_t bytes = (uint32_t)(end - cell_ - 8);
        os_.seekp(cell_ + 4);
        os_.write(reinterpret_cast<const char*>(&bytes), 4);
        os_.seekp(end);
        cell_ = -1;
    }

    // A 3-D double array, element (i, j, k) at data[i + ni*(j + nj*k)].
    void addArray(const std::string& name, const double* data, int ni, int nj, int nk)
    {
        const uint64_t n = (uint64_t)ni * nj * nk;
        std::vector<char> b;
        put_u32(b, kMiMatrix);
        put_u32(b, 0);
        put_flags(b, kMxDouble);
        const int32_t dims[3] = {ni, nj, nk};
        put_dims(b, dims, 3);
        put_name(b, name);
        put_u32(b, kMiDouble);
        put_u32(b, (uint32_t)(n * 8));
        const uint32_t bytes = (uint32_t)(b.size() - 8 + n * 8);
        std::memcpy(b.data() + 4, &bytes, 4);
        os_.write(b.data(), (std::streamsize)b.size());
        os_.write(reinterpret_cast<const char*>(data), (std::streamsize)(n * 8));
        if (cell_ >= 0) ++cell_added_;
    }

    // A field type with calculationDomain(), operator()(i, j, k) on a host copy and
    // SynchronizeHostStorage() (the reference's MatFile::addField, src/MatFile.h:63-168):
    // the interior is written, named `name` or `name_inc` when inc >= 0.
    template <typename Field>
    void addField(const std::string& name, const Field& f, int inc = -1)
    {
        f.SynchronizeHostStorage();
        const auto& d = f.calculationDomain();
        const int ni = d.iSize(), nj = d.jSize(), nk = d.kSize();
        std::vector<double> a((size_t)ni * nj * nk);
        for (int k = 0; k < nk; ++k)
            for (int j = 0; j < nj; ++j)
                for (int i = 0; i < ni; ++i) a[i + (size_t)ni * (j + (size_t)nj * k)] = f(i, j, k);
        addArray(inc >= 0 ? name + "_" + std::to_string(inc) : name, a.data(), ni, nj, nk);
    }
    template <typename Field>
    void addField(const Field& f, int inc = -1)
    {
        addField(f.name(), f, inc);
    }

    void close()
    {
        if (!os_.is_open()) return;
        endCell();
        os_.close();
    }

private:
    static constexpr uint32_t kMiInt8 = 1, kMiInt32 = 5, kMiUInt32 = 6, kMiDouble = 9, kMiMatrix = 14;
    static constexpr uint32_t kMxCell = 1, kMxDouble = 6;

    static void put_u32(std::vector<char>& b, uint32_t v)
    {
        const char* p = reinterpret_cast<const char*>(&v);
        b.insert(b.end(), p, p + 4);
    }
    static void pad8(std::vector<char>& b)
    {
        while (b.size() % 8) b.push_back(0);
    }
    static void put_flags(std::vector<char>& b, uint32_t cls)
    {
        put_u32(b, kMiUInt32);
        put_u32(b, 8);
        put_u32(b, cls);
        put_u32(b, 0);
    }
    static void put_dims(std::vector<char>& b, const int32_t* dims, int nd)
    {
        put_u32(b, kMiInt32);
        put_u32(b, (uint32_t)(4 * nd));
        for (int d = 0; d < nd; ++d) put_u32(b, (uint32_t)dims[d]);
        pad8(b);
    }
    static void put_name(std::vector<char>& b, const std::string& name)
    {
        put_u32(b, kMiInt8);
        put_u32(b, (uint32_t)name.size());
        b.insert(b.end(), name.begin(), name.end());
        pad8(b);
    }

    std::ofstream os_;
    long long cell_ = -1;
    int cell_n_ = 0, cell_added_ = 0;
};

}  // namespace sp
