// Logging helper kept for source compatibility (include/dwarves_debug.h).
#ifndef DWARVES_DEBUG_H
#define DWARVES_DEBUG_H

#include <iostream>

class Debug {
   public:
    static Debug* get_instance() {
        static Debug d;
        return &d;
    }
    template <typename... Args>
    void log(Args&&... args) {
        (void)sizeof...(args);
    }
    template <typename... Args>
    void print(Args&&... args) {
        ((std::cout << args), ...);
        std::cout << std::endl;
    }
    void enter_function(const char*) {}
    void leave_function(const char*) {}
};

#endif
