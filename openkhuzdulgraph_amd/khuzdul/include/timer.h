// Wall clock used by the apps' timing loops (include/timer.h:9 of the reference).
#ifndef TIMER_H
#define TIMER_H

double get_time();

#endif
