// Test driver for include/tsm_traj.hpp: minimal Point/Traj types shaped like
// the reference's (query/cpp/spatial/point.h, traj.h), trajectory pairs read
// from a text file, every measure computed through tsm_traj_simi, results
// printed one per line.
//
// Input: first line "n"; then 2n trajectories (the first n are trajs1), each
// as a line "len x0 y0 x1 y1 ...".  Output: for each measure in argv[2..],
// n lines "measure i value".  Exit status 2 with the library's message when a
// call fails (for example without a GPU).
#include <cstdio>
#include <fstream>
#include <iostream>
#include <memory>
#include <string>
#include <vector>

#include "tsm_traj.hpp"

struct Point {
    float x, y;
};
struct Traj {
    std::vector<Point *> points;
    size_t NumPoints() const { return points.size(); }
    ~Traj() {
        for (Point *p : points) delete p;
    }
};

int main(int argc, char **argv) {
    if (argc < 3) {
        std::cerr << "usage: traj_bridge FILE MEASURE...\n";
        return 1;
    }
    std::ifstream in(argv[1]);
    int n = 0;
    in >> n;
    std::vector<std::unique_ptr<Traj>> own;
    std::vector<Traj *> t1, t2;
    for (int i = 0; i < 2 * n; ++i) {
        int len = 0;
        in >> len;
        auto t = std::make_unique<Traj>();
        for (int j = 0; j < len; ++j) {
            float x, y;
            in >> x >> y;
            t->points.push_back(new Point{x, y});
        }
        (i < n ? t1 : t2).push_back(t.get());
        own.push_back(std::move(t));
    }
    for (int a = 2; a < argc; ++a) {
        const std::string m = argv[a];
        std::vector<float> d;
        const int rc = tsm_traj_simi(t1, t2, m, d, 0.0025);
        if (rc != TSM_OK) {
            std::cerr << "tsm_traj_simi(" << m << ") failed: " << rc << " " << tsm_last_error() << "\n";
            return 2;
        }
        for (size_t i = 0; i < d.size(); ++i) std::printf("%s %zu %.9g\n", m.c_str(), i, (double)d[i]);
    }
    return 0;
}
