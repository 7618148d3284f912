// PreIntegrationTable subset
// (include/voreen/core/datastructures/transfunc/1d/preintegrationtable.h):
// an N x N RGBA float table, element [fe*N + fs].
#pragma once
#include <vector>

#include "tgt/vector.h"

namespace voreen {

class PreIntegrationTable {
public:
    PreIntegrationTable(size_t n, std::vector<float>&& table) : n_(n), table_(std::move(table)) {}
    const tgt::vec4* getTable() const { return reinterpret_cast<const tgt::vec4*>(table_.data()); }
    size_t getDimension() const { return n_; }

private:
    size_t n_;
    std::vector<float> table_;
};

}  // namespace voreen
