// app_common.hpp -- command line and error handling shared by the application mains.
//
// Usage as the reference apps (dm-heom/src/heom/command_line.cpp:19-40):
//   <app> [--help] [--no-progress] <ocl_config_filename> <heom_config_filename>
// The first positional file configured the OpenCL device in the reference; it is
// accepted so that existing job scripts keep working, must exist, and is otherwise
// ignored (the device is the current HIP device).  Errors print a message and exit
// with -1 as dm-heom/src/heom/handle_main_exception.cpp:15-33 does.
#pragma once

#include <cstdio>
#include <fstream>
#include <stdexcept>
#include <string>
#include <vector>

namespace heom {

struct command_line {
    std::string ocl_config_filename;
    std::string heom_config_filename;
    bool no_progress = false;
    std::vector<std::pair<std::string, std::string>> options;  // extra --key value pairs

    command_line(int argc, char** argv, const std::vector<std::string>& value_options = {})
    {
        std::vector<std::string> pos;
        for (int a = 1; a < argc; ++a) {
            const std::string s = argv[a];
            if (s == "--help" || s == "-h")
                throw std::runtime_error(usage(argv[0], value_options) + "Help option was specified, just showing usage.");
            if (s == "--no-progress") {
                no_progress = true;
                continue;
            }
            bool taken = false;
            for (auto& o : value_options)
                if (s == "--" + o) {
                    if (a + 1 >= argc)
                        throw std::runtime_error(usage(argv[0], value_options) + "option --" + o + " needs a value");
                    options.emplace_back(o, argv[++a]);
                    taken = true;
                }
            if (!taken)
                pos.push_back(s);
        }
        if (pos.size() != 2)
            throw std::runtime_error(usage(argv[0], value_options) +
                                     "expected <ocl_config_filename> <heom_config_filename>");
        ocl_config_filename = pos[0];
        heom_config_filename = pos[1];
        std::ifstream f(ocl_config_filename);
        if (f.fail())
            throw std::runtime_error("could not open OpenCL config file: " + ocl_config_filename);
    }

    static std::string usage(const char* app, const std::vector<std::string>& value_options)
    {
        std::string u = std::string("\nUsage: ") + app + " [--no-progress]";
        for (auto& o : value_options)
            u += " [--" + o + " <value>]";
        return u + " <ocl_config_filename> <heom_config_filename>\n\n";
    }

    const std::string* option(const std::string& key) const
    {
        for (auto& kv : options)
            if (kv.first == key)
                return &kv.second;
        return nullptr;
    }
};

inline int report_exception(const std::exception& e)
{
    std::fprintf(stderr, "heom: error: %s\n", e.what());
    return -1;
}

}  // namespace heom
