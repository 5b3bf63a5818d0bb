// Host TensorMap returned through the C API (metatensor-core compatible subset).
#pragma once

#include <algorithm>
#include <cstdint>
#include <memory>
#include <string>
#include <vector>

#include "../../include/metatensor.h"

namespace rascaline_amd {

struct LabelsData {
    std::vector<std::string> names;
    std::vector<const char*> name_ptrs;
    std::vector<int32_t> values;  // count x size

    LabelsData() = default;
    LabelsData(std::vector<std::string> n, std::vector<int32_t> v) : names(std::move(n)), values(std::move(v)) { finish(); }
    // name_ptrs point into `names`: rebuild them on every copy / move
    LabelsData(const LabelsData& o) : names(o.names), values(o.values) { finish(); }
    LabelsData(LabelsData&& o) noexcept : names(std::move(o.names)), values(std::move(o.values)) { finish(); }
    LabelsData& operator=(const LabelsData& o) {
        names = o.names;
        values = o.values;
        finish();
        return *this;
    }
    LabelsData& operator=(LabelsData&& o) noexcept {
        names = std::move(o.names);
        values = std::move(o.values);
        finish();
        return *this;
    }
    void finish() {
        name_ptrs.clear();
        for (auto& s : names) name_ptrs.push_back(s.c_str());
    }
    size_t size() const { return names.size(); }
    size_t count() const { return names.empty() ? 0 : values.size() / names.size(); }
    void fill(mts_labels_t* out) const;

    static LabelsData range(const std::string& name, int n) {
        std::vector<int32_t> v(n);
        for (int i = 0; i < n; i++) v[i] = i;
        return LabelsData({name}, v);
    }
};

// float64 storage of a block's array. allocate() leaves the elements
// uninitialised: the producers of a block's data write every element, and value-
// initialising arrays of tens of GB would cost as much as filling them.
class DoubleBuffer {
public:
    DoubleBuffer() = default;
    DoubleBuffer(DoubleBuffer&&) noexcept = default;
    DoubleBuffer& operator=(DoubleBuffer&&) noexcept = default;
    DoubleBuffer& operator=(const std::vector<double>& v) {
        allocate(v.size());
        std::copy(v.begin(), v.end(), ptr_.get());
        return *this;
    }
    void allocate(size_t n) {
        ptr_.reset(n ? new double[n] : nullptr);
        size_ = n;
    }
    void assign(size_t n, double value) {
        allocate(n);
        std::fill_n(ptr_.get(), n, value);
    }
    double* data() { return ptr_.get(); }
    const double* data() const { return ptr_.get(); }
    size_t size() const { return size_; }
    double& operator[](size_t k) { return ptr_[k]; }
    const double& operator[](size_t k) const { return ptr_[k]; }
    double* begin() { return ptr_.get(); }
    double* end() { return ptr_.get() + size_; }

private:
    std::unique_ptr<double[]> ptr_;
    size_t size_ = 0;
};

// Contiguous float64 array owned by a block.
struct ArrayData {
    std::vector<uintptr_t> shape;
    DoubleBuffer data;
};

}  // namespace rascaline_amd

struct mts_block_t {
    rascaline_amd::LabelsData samples;
    std::vector<rascaline_amd::LabelsData> components;
    rascaline_amd::LabelsData properties;
    rascaline_amd::ArrayData array;
    std::vector<std::string> gradient_names;
    std::vector<const char*> gradient_name_ptrs;
    std::vector<std::unique_ptr<mts_block_t>> gradients;

    void add_gradient(const std::string& name, std::unique_ptr<mts_block_t> g) {
        gradient_names.push_back(name);
        gradients.push_back(std::move(g));
        gradient_name_ptrs.clear();
        for (auto& s : gradient_names) gradient_name_ptrs.push_back(s.c_str());
    }
};

struct mts_tensormap_t {
    rascaline_amd::LabelsData keys;
    std::vector<std::unique_ptr<mts_block_t>> blocks;
};
