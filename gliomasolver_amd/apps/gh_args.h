// Command-line options with the reference's conventions
// (MRAG/MRAGio/MRAG_IO_ArgumentParser.h, tools/pi4u_lite/Inference/Likelihood/
// source/ArgumentParser.h): a word starting with '-' is a key, the words up to
// the next key are its value (joined by spaces), a key with no value reads as
// "1"; an absent key yields the default; asBool is false for "0" and "false".
#pragma once
#include <cstdlib>
#include <map>
#include <string>

namespace gliomahip {

class Args {
public:
    Args(int argc, const char** argv) {
        for (int i = 1; i < argc; ++i) {
            if (argv[i][0] != '-') continue;
            std::string v;
            int j = i + 1;
            for (; j < argc && argv[j][0] != '-'; ++j) {
                if (!v.empty()) v += ' ';
                v += argv[j];
            }
            if (j == i + 1) v = "1";
            kv_[argv[i]] = v;
            i = j - 1;
        }
    }
    bool has(const std::string& k) const { return kv_.count(k) && !kv_.at(k).empty(); }
    std::string str(const std::string& k, const std::string& def = "") const { return has(k) ? kv_.at(k) : def; }
    double num(const std::string& k, double def = 0) const { return has(k) ? std::atof(kv_.at(k).c_str()) : def; }
    int integer(const std::string& k, int def = 0) const { return has(k) ? std::atoi(kv_.at(k).c_str()) : def; }
    bool flag(const std::string& k, bool def = false) const {
        if (!has(k)) return def;
        const std::string& v = kv_.at(k);
        return !(v == "0" || v == "false");
    }

private:
    std::map<std::string, std::string> kv_;
};

}  // namespace gliomahip
