// Input-file front end of the reference's drivers (code/modules/read_infile.cpp:16-173,
// code/include/read_infile.hpp): the YAML schema of code/builds/test.yaml, check_and_assign_value,
// params_class with its output-file checks, and read_params for "main -i infile".
//
// The image has no yaml-cpp, so this header carries a reader for the subset of YAML those inputs use:
// block mappings by indentation, "key: scalar" and "key:" + nested block, '#' comments, plain or
// quoted scalars.  Errors are std::runtime_error with the reference's messages (it calls
// Kokkos::abort).  Two deliberate differences, both fixes of reference pitfalls:
//   * int keys are read like the reference, through stod (read_infile.cpp:47), so "1e3" is 1000;
//   * a missing randomize_trajectory means 0 (HMC.cpp:35 requires it, yet test.yaml lacks it).
#pragma once

#include <cstdio>
#include <fstream>
#include <iostream>
#include <memory>
#include <sstream>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

namespace chmc_host {

class yaml_node {
public:
    yaml_node() = default;
    explicit yaml_node(std::string v) : scalar_(std::move(v)), defined_(true) {}

    explicit operator bool() const { return defined_; }
    bool is_map() const { return is_map_; }
    const std::string& scalar() const { return scalar_; }

    const yaml_node& operator[](const std::string& key) const
    {
        static const yaml_node undefined;
        for (const auto& kv : map_)
            if (kv.first == key) return *kv.second;
        return undefined;
    }

    template <class T> T as() const;

    yaml_node& add_child(const std::string& key)
    {
        is_map_ = true;
        defined_ = true;
        map_.emplace_back(key, std::make_unique<yaml_node>());
        map_.back().second->defined_ = true;
        return *map_.back().second;
    }
    void set_scalar(const std::string& v) { scalar_ = v; defined_ = true; }

    void print(std::ostream& os, int indent = 0) const
    {
        for (const auto& kv : map_) {
            os << std::string(indent, ' ') << kv.first << ":";
            if (kv.second->is_map_) { os << "\n"; kv.second->print(os, indent + 2); }
            else os << " " << kv.second->scalar_ << "\n";
        }
    }

private:
    std::string scalar_;
    bool defined_ = false, is_map_ = false;
    std::vector<std::pair<std::string, std::unique_ptr<yaml_node>>> map_;
};

template <> inline std::string yaml_node::as<std::string>() const
{
    if (is_map_) throw std::runtime_error("not a scalar");
    return scalar_;
}
template <> inline double yaml_node::as<double>() const
{
    size_t pos = 0;
    const double v = std::stod(as<std::string>(), &pos);
    if (pos != scalar_.size()) throw std::runtime_error("not a number");
    return v;
}
template <> inline int yaml_node::as<int>() const { return (int)std::stod(as<std::string>()); }   // read_infile.cpp:47
template <> inline bool yaml_node::as<bool>() const
{
    const std::string s = as<std::string>();
    if (s == "true" || s == "True" || s == "TRUE" || s == "yes" || s == "on") return true;
    if (s == "false" || s == "False" || s == "FALSE" || s == "no" || s == "off") return false;
    throw std::runtime_error("not a boolean");
}

inline std::string yaml_trim(const std::string& s)
{
    const size_t b = s.find_first_not_of(" \t\r");
    if (b == std::string::npos) return "";
    const size_t e = s.find_last_not_of(" \t\r");
    return s.substr(b, e - b + 1);
}

// the block-mapping subset of YAML used by the reference's input files
inline yaml_node yaml_parse(std::istream& in, const std::string& name = "<input>")
{
    yaml_node doc;
    struct level { int indent; yaml_node* node; };
    std::vector<level> stack{ { -1, &doc } };
    std::string line;
    int lineno = 0;
    yaml_node* pending = nullptr;     // "key:" awaiting a nested block
    int pending_indent = -1;
    while (std::getline(in, line)) {
        lineno++;
        // strip comments outside quotes
        bool sq = false, dq = false;
        for (size_t i = 0; i < line.size(); i++) {
            const char ch = line[i];
            if (ch == '\'' && !dq) sq = !sq;
            else if (ch == '"' && !sq) dq = !dq;
            else if (ch == '#' && !sq && !dq && (i == 0 || line[i - 1] == ' ' || line[i - 1] == '\t')) { line.resize(i); break; }
        }
        if (yaml_trim(line).empty() || yaml_trim(line) == "---") continue;
        if (line.find('\t') != std::string::npos && line.find_first_not_of(" \t") > line.find('\t'))
            throw std::runtime_error(name + ":" + std::to_string(lineno) + ": tabs are not allowed for indentation");
        const int indent = (int)line.find_first_not_of(' ');
        const std::string body = yaml_trim(line);
        const size_t colon = body.find(':');
        if (colon == std::string::npos || body[0] == '-')
            throw std::runtime_error(name + ":" + std::to_string(lineno) + ": expected 'key: value'");
        const std::string key = yaml_trim(body.substr(0, colon));
        std::string val = yaml_trim(body.substr(colon + 1));
        if (pending) {
            if (indent > pending_indent) stack.push_back({ pending_indent, pending });
            pending = nullptr;
        }
        while (stack.size() > 1 && indent <= stack.back().indent) stack.pop_back();
        yaml_node& child = stack.back().node->add_child(key);
        if (val.empty()) {
            pending = &child;
            pending_indent = indent;
        } else {
            if (val.size() >= 2 && ((val.front() == '"' && val.back() == '"') || (val.front() == '\'' && val.back() == '\'')))
                val = val.substr(1, val.size() - 2);
            child.set_scalar(val);
        }
    }
    return doc;
}

inline yaml_node yaml_load_file(const std::string& path)
{
    std::ifstream f(path);
    if (!f.good()) throw std::runtime_error("unable to open input file " + path);
    return yaml_parse(f, path);
}

// check_and_assign_value (read_infile.cpp:16-53)
template <class T> T check_and_assign_value(const yaml_node& doc, const char* tag)
{
    if (!doc[tag]) {
        std::printf("error: tag %s is not present in the input file\n", tag);
        std::cout << "infile structure:" << std::endl;
        doc.print(std::cout);
        throw std::runtime_error("params not found");
    }
    try {
        return doc[tag].template as<T>();
    } catch (const std::exception&) {
        std::printf("error: impossible to read tag %s\n", tag);
        throw std::runtime_error("Incorrect input type");
    }
}

inline bool file_exist(const std::string& name)
{
    std::ifstream f(name.c_str());
    return f.good();
}

// params_class (read_infile.hpp:7-32, read_infile.cpp:89-149)
struct params_class {
    double L[3] = { 0, 0, 0 };
    bool append = false;
    int seed = 0;
    int istart = 0;
    FILE* fileout = nullptr;
    std::string StartCondition, start_configuration_file, nameout, rng_host_state, rng_device_state, name_RDF;

    params_class() = default;
    // touch_files = false parses and validates the values only (no existence checks, no output opened)
    params_class(const yaml_node& doc, bool check_overwrite = true, bool touch_files = true)
    {
        L[0] = check_and_assign_value<double>(doc["geometry"], "Lx");
        L[1] = check_and_assign_value<double>(doc["geometry"], "Ly");
        L[2] = check_and_assign_value<double>(doc["geometry"], "Lz");
        seed = check_and_assign_value<int>(doc, "seed");
        StartCondition = check_and_assign_value<std::string>(doc, "StartCondition");
        std::printf("geometry:\n   Lx: %g\n   Ly: %g\n   Lz: %g\nseed: %d\nStartCondition: %s\n", L[0], L[1], L[2], seed,
                    StartCondition.c_str());
        if (StartCondition == "read")
            start_configuration_file = check_and_assign_value<std::string>(doc, "start_configuration_file");
        nameout = check_and_assign_value<std::string>(doc, "output_file");
        rng_host_state = check_and_assign_value<std::string>(doc, "rng_host_state");
        rng_device_state = check_and_assign_value<std::string>(doc, "rng_device_state");
        if (rng_device_state == rng_host_state) throw std::runtime_error("rng_device_state must be different from rng_host_state");
        if (rng_device_state == nameout) throw std::runtime_error("rng_device_state must be different from output_file");
        if (rng_host_state == nameout) throw std::runtime_error("rng_host_state must be different from output_file");
        append = check_and_assign_value<bool>(doc, "append");
        if (append) {
            if (StartCondition != "read")
                throw std::runtime_error("append=true so the start condition must be read, while in the input file StartCondition=" +
                                         StartCondition);
            if (start_configuration_file != nameout)
                throw std::runtime_error("append=true is only supported if the start_configuration_file is the same as the output_file");
            if (touch_files)
                for (const std::string& f : { rng_host_state, rng_device_state })
                    if (!file_exist(f)) throw std::runtime_error("unable to open file " + f);
        } else if (check_overwrite && touch_files) {
            for (const std::string& f : { nameout, rng_host_state, rng_device_state })
                if (file_exist(f))
                    throw std::runtime_error("output file " + f + " exists but the input file has append=false");
        }
        if (touch_files) {
            if (nameout.empty() || nameout == "null") throw std::runtime_error("unable to open file " + nameout);
            fileout = std::fopen(nameout.c_str(), "ab");
            if (!fileout) throw std::runtime_error("unable to open file " + nameout);
        }
        istart = 0;
        if (doc["particles"]["RDF"]) name_RDF = check_and_assign_value<std::string>(doc["particles"]["RDF"], "output_file");
    }
};

// read_params (read_infile.cpp:152-172): the file after "-i"
inline yaml_node read_params(int argc, char** argv)
{
    int opt = -1;
    for (int i = 1; i < argc; ++i)
        if (std::string(argv[i]) == "-i") { opt = i + 1; break; }
    if (opt < 0 || opt == argc) {
        std::cout << "No input file specified, Aborting" << std::endl;
        std::cout << "usage:  ./main -i infile.in" << std::endl;
        throw std::runtime_error("no input file");
    }
    std::cout << "Trying input file " << argv[opt] << std::endl;
    return yaml_load_file(argv[opt]);
}

}  // namespace chmc_host
