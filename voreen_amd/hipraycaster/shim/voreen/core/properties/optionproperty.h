// OptionProperty subset (include/voreen/core/properties/optionproperty.h).
#pragma once
#include "voreen/core/shim_base.h"

namespace voreen {

template <typename T>
class OptionProperty : public Property {
public:
    OptionProperty(const std::string& id, const std::string& gui, int inv = Processor::INVALID_RESULT)
        : Property(id, gui, inv) {}
    void addOption(const std::string& key, const std::string& desc, const T& value) {
        keys_.push_back(key);
        descs_.push_back(desc);
        values_.push_back(value);
        if (sel_ < 0) sel_ = 0;
    }
    void selectByKey(const std::string& key) {
        for (size_t i = 0; i < keys_.size(); ++i)
            if (keys_[i] == key) { sel_ = (int)i; changed(); return; }
    }
    void select(const std::string& key) { selectByKey(key); }
    const std::string& getKey() const { static const std::string none; return sel_ < 0 ? none : keys_[sel_]; }
    T getValue() const { return sel_ < 0 ? T() : values_[sel_]; }
    bool hasKey(const std::string& k) const { for (auto& x : keys_) if (x == k) return true; return false; }
    std::string str() const override { return getKey(); }
    bool set(const std::string& k) override { if (!hasKey(k)) return false; selectByKey(k); return true; }
    std::vector<std::string> optionKeys() const override { return keys_; }

protected:
    std::vector<std::string> keys_, descs_;
    std::vector<T> values_;
    int sel_ = -1;
};

class IntOptionProperty : public OptionProperty<int> {
public:
    IntOptionProperty(const std::string& id, const std::string& gui, int inv = Processor::INVALID_RESULT)
        : OptionProperty<int>(id, gui, inv) {}
};

class StringOptionProperty : public OptionProperty<std::string> {
public:
    StringOptionProperty(const std::string& id, const std::string& gui, int inv = Processor::INVALID_RESULT)
        : OptionProperty<std::string>(id, gui, inv) {}
    using OptionProperty<std::string>::addOption;
    void addOption(const std::string& key, const std::string& desc) { OptionProperty<std::string>::addOption(key, desc, key); }
};

}  // namespace voreen
