// rt_ascii.h -- the line reader of the reference's text inputs (range files,
// CCD specifications): AsciiFile::ReadLine (common-lib/ascii_file.cpp:88-128).
// A line is data when it starts (after leading spaces) with n numbers, a
// comment when it starts with '#', empty when it holds only spaces, invalid
// otherwise; a last line without a newline is not read (getline, then eof(),
// :99-100).
#ifndef RT_ASCII_H
#define RT_ASCII_H

#include <stdlib.h>

#include <fstream>
#include <string>
#include <utility>
#include <vector>

namespace rtascii {

enum LineFlag { kData = 0, kInvalid = 1, kComment = 2, kEmpty = 3 };

class Reader {
public:
    explicit Reader(const std::string& path) : in(path) {}
    bool ok() const { return (bool)in; }

    // next line as (flag, numbers); false at the end of the file
    bool next(size_t n_elems, int& flag, std::vector<double>& v) {
        std::string line;
        if (!std::getline(in, line) || in.eof()) return false;
        v.clear();
        size_t i = 0;
        while (i < line.size() && line[i] == ' ') ++i;
        if (i == line.size()) {
            flag = kEmpty;
            return true;
        }
        const char* s = line.c_str() + i;
        if (s[0] == '#') {
            flag = kComment;
            return true;
        }
        size_t k = 0;
        for (; k < n_elems; ++k) {
            char* e = nullptr;
            const double x = strtod(s, &e);
            if (e == s) break;
            v.push_back(x);
            s = e;
        }
        flag = k == n_elems ? kData : kInvalid;
        return true;
    }

private:
    std::ifstream in;
};

}  // namespace rtascii

#endif  // RT_ASCII_H
