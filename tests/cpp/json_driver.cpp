// Command-line front end of the JSON layer (the reference's examples/rcwa_json.cpp): reads one request from
// a file (or stdin with "-"), writes the reply to stdout.  --indent N writes it over several lines.
// Exit status 0 when the reply has ok = true, 1 otherwise.
#include <fstream>
#include <iostream>
#include <iterator>
#include <sstream>
#include <string>

#include "stablercwa_amd/rcwa_json.hpp"

int main(int argc, char** argv) {
    int indent = -1;
    std::string path = "-";
    for (int i = 1; i < argc; ++i) {
        const std::string a = argv[i];
        if (a == "--indent" && i + 1 < argc) indent = std::stoi(argv[++i]);
        else path = a;
    }
    std::string text;
    if (path == "-") {
        text.assign(std::istreambuf_iterator<char>(std::cin), std::istreambuf_iterator<char>());
    } else {
        std::ifstream f(path, std::ios::binary);
        if (!f) {
            std::cerr << "cannot read " << path << "\n";
            return 2;
        }
        text.assign(std::istreambuf_iterator<char>(f), std::istreambuf_iterator<char>());
    }
    const srcwa_amd::rcwa::JsonReply r = srcwa_amd::rcwa::solve_json_reply(text, indent);
    std::cout << r.text << "\n";
    return r.ok ? 0 : 1;
}
