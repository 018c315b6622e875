// rts::arglist (reference rts/tools/arguments.h): "--name v1 v2 ..." options
// with defaults.  An option with a non-empty default counts as set; values
// are split on single spaces; as_float / as_int read a value or give 0.
#pragma once

#include <iomanip>
#include <iostream>
#include <sstream>
#include <stdexcept>
#include <string>
#include <vector>

namespace rts {

class argument {
   public:
    argument(std::string name_, std::string desc, std::string def = "", std::string range_ = "")
        : name(std::move(name_)), range(std::move(range_)) {
        split(desc, '\n', &descr);
        set_values(def);
        flag = !def.empty();
    }

    int nargs() const { return (int)vals.size(); }
    bool is_set() const { return flag; }
    const std::string& id() const { return name; }

    std::string as_text(int n = 0) const {
        require();
        return n < (int)vals.size() ? vals[n] : std::string();
    }
    float as_float(int n = 0) const {
        require();
        if (n >= (int)vals.size()) return 0;
        float r;
        std::istringstream is(vals[n]);
        if (!(is >> r)) r = 0;
        return r;
    }
    int as_int(int n = 0) const {
        require();
        if (n >= (int)vals.size()) return 0;
        int r;
        std::istringstream is(vals[n]);
        if (!(is >> r)) r = 0;
        return r;
    }

    void set(const std::string& v) {
        set_values(v);
        flag = true;
    }

    int col_width() const {
        int n = 3 + (int)name.size();
        if (!vals.empty()) {
            n += 6;
            for (const std::string& v : vals) n += (int)v.size() + 1;
        }
        return n + 4;
    }

    std::string toStr(int width) const {
        std::stringstream ss;
        std::string left = " --" + name;
        if (!vals.empty()) {
            left += " ( = ";
            for (const std::string& v : vals) left += v + " ";
            left += ")";
        }
        ss << std::left << std::setw(width) << left;
        for (size_t d = 0; d < descr.size(); ++d) {
            if (d == 0) ss << descr[0];
            else ss << std::endl << std::setfill(' ') << std::setw(width) << " " << descr[d];
        }
        if (!range.empty()) ss << std::endl << std::setfill(' ') << std::setw(width) << " " << "    " << range;
        return ss.str();
    }

   private:
    std::string name;
    std::vector<std::string> descr, vals;
    std::string range;
    bool flag = false;

    static void split(const std::string& s, char sep, std::vector<std::string>* out) {
        out->clear();
        std::stringstream ss(s);
        std::string item;
        while (std::getline(ss, item, sep)) out->push_back(item);
    }
    void set_values(const std::string& v) { split(v, ' ', &vals); }
    void require() const {
        if (!flag) throw std::runtime_error("ERROR - Argument requested without being set: " + name);
    }
};

class arglist {
   public:
    void section(const std::string& name) { sections.push_back({name, args.size()}); }

    void add(std::string name, std::string desc, std::string def = "", std::string range = "") {
        args.emplace_back(name, desc, def, range);
        width = std::max(width, args.back().col_width());
    }

    // Unknown option names are reported and ignored, as in the reference.
    void parse(int argc, const char* const* argv) {
        if (argc <= 1) return;
        std::string name, params;
        for (int i = 1; i < argc; ++i) {
            if (argv[i][0] == '-' && argv[i][1] == '-') {
                if (!name.empty()) set(name, params);
                name = argv[i] + 2;
                params.clear();
            } else {
                if (!params.empty()) params += " ";
                params += argv[i];
            }
        }
        set(name, params);
    }

    void set(const std::string& name, const std::string& value) {
        argument* a = find(name);
        if (a) {
            a->set(value);
            width = std::max(width, a->col_width());
        } else {
            std::cout << "ERROR - Argument not recognized: " << name << std::endl;
        }
    }

    bool operator()(const std::string& name) const { return get(name).is_set(); }
    const argument& operator[](const std::string& name) const { return get(name); }

    std::string toStr() const {
        std::stringstream ss;
        size_t si = 0;
        for (size_t a = 0; a < args.size(); ++a) {
            while (si < sections.size() && sections[si].second == a) {
                ss << std::endl << std::left << std::setfill('=') << std::setw(width) << sections[si].first
                   << std::setfill(' ') << std::endl;
                ++si;
            }
            ss << args[a].toStr(width) << std::endl;
        }
        return ss.str();
    }

   private:
    std::vector<argument> args;
    std::vector<std::pair<std::string, size_t>> sections;
    int width = 0;

    argument* find(const std::string& name) {
        for (argument& a : args)
            if (a.id() == name) return &a;
        return nullptr;
    }
    const argument& get(const std::string& name) const {
        for (const argument& a : args)
            if (a.id() == name) return a;
        throw std::runtime_error("ERROR - Unspecified parameter name: " + name);
    }
};

}  // namespace rts
