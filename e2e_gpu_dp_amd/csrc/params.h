// params.h — runtime reader of the builder's parameter file.
//
// The reference compiles its parameters into the builder: src/input_to_build_model.h is
// #included inside main (src/build_model.cu:932). Here the same file is read at run time.
// It accepts that header's own syntax — C declarations `type name = expr;` with
// arithmetic over numbers and earlier names, C casts, integer division between ints,
// string literals, // and /* */ comments — and plain `name = value` lines.
#pragma once

#include <cctype>
#include <cmath>
#include <fstream>
#include <map>
#include <sstream>
#include <stdexcept>
#include <string>

namespace params {

struct Value {
    double v = 0.0;
    bool is_int = false;
    bool is_str = false;
    std::string s;
};

class Parser {
   public:
    std::map<std::string, Value> vals;

    void parse_text(const std::string& text0) {
        std::string text = strip_comments(text0);
        std::string stmt;
        for (char c : text) {
            if (c == ';' || c == '\n') {
                statement(stmt);
                stmt.clear();
            } else {
                stmt += c;
            }
        }
        statement(stmt);
    }

    void parse_file(const std::string& path) {
        std::ifstream f(path);
        if (!f) throw std::runtime_error("cannot open parameter file " + path);
        std::stringstream ss;
        ss << f.rdbuf();
        parse_text(ss.str());
    }

    bool has(const std::string& k) const { return vals.count(k) > 0; }
    double num(const std::string& k) const {
        auto it = vals.find(k);
        if (it == vals.end() || it->second.is_str) throw std::runtime_error("parameter '" + k + "' missing");
        return it->second.v;
    }
    double num(const std::string& k, double dflt) const { return has(k) ? num(k) : dflt; }
    std::string str(const std::string& k) const {
        auto it = vals.find(k);
        if (it == vals.end() || !it->second.is_str) throw std::runtime_error("parameter '" + k + "' missing");
        return it->second.s;
    }

   private:
    std::string src_;
    size_t pos_ = 0;

    static std::string strip_comments(const std::string& t) {
        std::string o;
        for (size_t i = 0; i < t.size(); i++) {
            if (t[i] == '"') {  // keep string literals intact
                size_t j = t.find('"', i + 1);
                if (j == std::string::npos) j = t.size() - 1;
                o += t.substr(i, j - i + 1);
                i = j;
            } else if (t.compare(i, 2, "//") == 0 || t[i] == '#') {
                while (i < t.size() && t[i] != '\n') i++;
                o += '\n';
            } else if (t.compare(i, 2, "/*") == 0) {
                size_t j = t.find("*/", i + 2);
                i = (j == std::string::npos) ? t.size() : j + 1;
            } else {
                o += t[i];
            }
        }
        return o;
    }

    static std::string trim(const std::string& s) {
        size_t b = s.find_first_not_of(" \t\r");
        if (b == std::string::npos) return "";
        size_t e = s.find_last_not_of(" \t\r");
        return s.substr(b, e - b + 1);
    }

    void statement(const std::string& s0) {
        std::string s = trim(s0);
        if (s.empty()) return;
        size_t eq = s.find('=');
        if (eq == std::string::npos) return;  // declarations without value (e.g. extern)
        std::string lhs = trim(s.substr(0, eq));
        std::string rhs = trim(s.substr(eq + 1));
        // lhs: [type words...] name
        std::string type, name;
        {
            std::istringstream ls(lhs);
            std::string w, last;
            while (ls >> w) {
                if (!last.empty()) type += (type.empty() ? "" : " ") + last;
                last = w;
            }
            name = last;
        }
        if (name.empty()) return;
        Value v;
        if (!rhs.empty() && rhs[0] == '"') {
            v.is_str = true;
            v.s = rhs.substr(1, rhs.rfind('"') - 1);
        } else {
            src_ = rhs;
            pos_ = 0;
            v = expr();
            skip();
            if (pos_ != src_.size()) throw std::runtime_error("cannot parse value of " + name + ": " + rhs);
            if (type.find("int") != std::string::npos || type.find("long") != std::string::npos) {
                v.v = std::trunc(v.v);
                v.is_int = true;
            } else if (type.find("float") != std::string::npos) {
                v.v = (double)(float)v.v;
                v.is_int = false;
            } else if (type.find("double") != std::string::npos) {
                v.is_int = false;
            }
        }
        vals[name] = v;
    }

    void skip() {
        while (pos_ < src_.size() && std::isspace((unsigned char)src_[pos_])) pos_++;
    }
    bool eat(char c) {
        skip();
        if (pos_ < src_.size() && src_[pos_] == c) {
            pos_++;
            return true;
        }
        return false;
    }
    static Value bin(const Value& a, const Value& b, char op) {
        Value r;
        r.is_int = a.is_int && b.is_int;
        switch (op) {
            case '+': r.v = a.v + b.v; break;
            case '-': r.v = a.v - b.v; break;
            case '*': r.v = a.v * b.v; break;
            default:
                if (b.v == 0) throw std::runtime_error("division by zero in parameter file");
                r.v = r.is_int ? std::trunc(a.v / b.v) : a.v / b.v;
        }
        return r;
    }
    Value expr() {
        Value a = term();
        for (;;) {
            if (eat('+')) a = bin(a, term(), '+');
            else if (eat('-')) a = bin(a, term(), '-');
            else return a;
        }
    }
    Value term() {
        Value a = unary();
        for (;;) {
            if (eat('*')) a = bin(a, unary(), '*');
            else if (eat('/')) a = bin(a, unary(), '/');
            else return a;
        }
    }
    Value unary() {
        if (eat('-')) {
            Value v = unary();
            v.v = -v.v;
            return v;
        }
        if (eat('+')) return unary();
        return primary();
    }
    Value primary() {
        skip();
        if (eat('(')) {
            // C cast?
            size_t save = pos_;
            std::string w;
            skip();
            while (pos_ < src_.size() && (std::isalnum((unsigned char)src_[pos_]) || src_[pos_] == '_'))
                w += src_[pos_++];
            if ((w == "int" || w == "int32_t" || w == "long" || w == "float" || w == "double") && eat(')')) {
                Value v = unary();
                if (w == "float") {
                    v.v = (double)(float)v.v;
                    v.is_int = false;
                } else if (w == "double") {
                    v.is_int = false;
                } else {
                    v.v = std::trunc(v.v);
                    v.is_int = true;
                }
                return v;
            }
            pos_ = save;
            Value v = expr();
            if (!eat(')')) throw std::runtime_error("missing ) in " + src_);
            return v;
        }
        skip();
        size_t b = pos_;
        if (pos_ < src_.size() && (std::isdigit((unsigned char)src_[pos_]) || src_[pos_] == '.')) {
            while (pos_ < src_.size() &&
                   (std::isalnum((unsigned char)src_[pos_]) || src_[pos_] == '.' ||
                    ((src_[pos_] == '-' || src_[pos_] == '+') && (src_[pos_ - 1] == 'e' || src_[pos_ - 1] == 'E'))))
                pos_++;
            std::string tok = src_.substr(b, pos_ - b);
            Value v;
            while (!tok.empty() && (tok.back() == 'f' || tok.back() == 'F' || tok.back() == 'L' || tok.back() == 'l'))
                tok.pop_back();
            v.v = std::stod(tok);
            v.is_int = tok.find_first_of(".eE") == std::string::npos;
            return v;
        }
        std::string w;
        while (pos_ < src_.size() && (std::isalnum((unsigned char)src_[pos_]) || src_[pos_] == '_'))
            w += src_[pos_++];
        if (w.empty()) throw std::runtime_error("cannot parse: " + src_);
        if (w == "M_PI") {
            Value v;
            v.v = M_PI;
            return v;
        }
        auto it = vals.find(w);
        if (it == vals.end() || it->second.is_str) throw std::runtime_error("unknown name '" + w + "' in " + src_);
        return it->second;
    }
};

}  // namespace params
