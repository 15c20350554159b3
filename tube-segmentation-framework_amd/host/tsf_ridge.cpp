// The ridge-traversal centerline method: runRidgeTraversal
// (ridgeTraversalCenterlineExtraction.cpp:37-369) and its driver
// runCircleFittingAndRidgeTraversal (tube-segmentation.cpp:1260-1350).
//
// The traversal is a sequential walk on the host, as in the reference: from
// start points taken in decreasing TDF, it steps to the neighbour ahead (a
// direction within ~84 degrees of the current one) with the largest
// 1 - |F| (or TDF * (1 - |F|) where the radius is below 1.5), re-aiming
// along the Hessian's tube direction at every step (doEigen, host code as
// in the reference), until it leaves the thresholds, meets an existing centerline or
// loops.  Two choices fix what the reference leaves open: the start points
// enter the priority queue in index order (the reference pushes them from an
// OpenMP loop), and among trees of equal length the smallest id is the
// largest (the reference takes whichever its hash map lists first).
#include <cmath>
#include <list>
#include <map>
#include <queue>
#include <stack>
#include <unordered_map>
#include <unordered_set>

#include "tsf_dropin.hpp"

namespace tsf {

namespace {

struct StartPoint {
    float value;
    int x, y, z;
};
struct ByValue {
    bool operator()(const StartPoint& a, const StartPoint& b) const { return a.value < b.value; }
};

float sign_of(float a) { return a < 0 ? -1.0f : 1.0f; }

float3 normalized(float3 v) {
    const float l = std::sqrt(v.x * v.x + v.y * v.y + v.z * v.z);
    return float3(v.x / l, v.y / l, v.z / l);
}
float dot3(const float3& a, const float3& b) { return a.x * b.x + a.y * b.y + a.z * b.z; }

}  // namespace

char* runRidgeTraversal(TubeSegmentation& T, int3 size, paramList& parameters,
                        std::stack<CenterlinePoint> centerlineStack) {
    const float Thigh = getParam(parameters, "tdf-high");
    const int Dmin = (int)getParam(parameters, "min-distance");
    const float Mlow = getParam(parameters, "m-low");
    const float Tlow = getParam(parameters, "tdf-low");
    const int maxBelowTlow = (int)getParam(parameters, "max-below-tdf-low");
    const float minMeanTube = getParam(parameters, "min-mean-tdf");
    const int TreeMin = (int)getParam(parameters, "min-tree-length");
    const int X = size.x, Y = size.y, Z = size.z;
    const size_t total = (size_t)X * Y * Z;
    auto at = [&](int x, int y, int z) { return (size_t)x + (size_t)X * ((size_t)y + (size_t)Y * z); };
    auto mag = [&](size_t i) {
        return std::sqrt(std::pow(T.Fx[i], 2.0f) + std::pow(T.Fy[i], 2.0f) + std::pow(T.Fz[i], 2.0f));
    };
    auto M = [&](size_t i) { return 1 - mag(i); };

    // start points: TDF >= tdf-high and |F| minimal in the 3x3x3 neighbourhood
    std::priority_queue<StartPoint, std::vector<StartPoint>, ByValue> queue;
    for (int z = 2; z < Z - 2; z++)
        for (int y = 2; y < Y - 2; y++)
            for (int x = 2; x < X - 2; x++) {
                const size_t i = at(x, y, z);
                if (T.TDF[i] < Thigh) continue;
                const float m0 = mag(i);
                bool valid = true;
                for (int c = -1; c < 2 && valid; c++)
                    for (int b = -1; b < 2 && valid; b++)
                        for (int a = -1; a < 2 && valid; a++)
                            if (mag(at(x + a, y + b, z + c)) < m0) valid = false;
                if (valid) queue.push(StartPoint{T.TDF[i], x, y, z});
            }
    if (queue.empty()) throw Error(TSF_ERR_RANGE, "no valid start points found");

    // voxel 0 becomes the "no candidate" reference: TDF 0, |F| = 1
    T.TDF[0] = 0;
    T.Fx[0] = 1;
    T.Fy[0] = 0;
    T.Fz[0] = 0;

    std::vector<int> centerlines(total, 0);
    int counter = 1;
    std::map<int, int> centerlineDistances;
    std::map<int, std::stack<CenterlinePoint>> centerlineStacks;

    while (!queue.empty()) {
        const StartPoint p = queue.top();
        queue.pop();
        if (centerlines[at(p.x, p.y, p.z)] == 1) continue;

        std::unordered_set<size_t> newCenterlines;
        newCenterlines.insert(at(p.x, p.y, p.z));
        int distance = 1, connections = 0, prevConnection = -1, secondConnection = -1;
        float meanTube = T.TDF[at(p.x, p.y, p.z)];
        std::stack<CenterlinePoint> stack;
        CenterlinePoint startPoint;
        startPoint.pos = int3(p.x, p.y, p.z);
        stack.push(startPoint);

        for (int direction = -1; direction < 3; direction += 2) {
            int belowTlow = 0;
            int3 position(p.x, p.y, p.z);
            float3 t_i = getTubeDirection(T, position, size);
            t_i = float3(t_i.x * direction, t_i.y * direction, t_i.z * direction);
            float3 t_i_1 = t_i;
            while (true) {
                int3 maxPoint(0, 0, 0);
                if (position.x < 3 || position.x > X - 3 || position.y < 3 || position.y > Y - 3 || position.z < 3 ||
                    position.z > Z - 3)
                    break;
                // the neighbour ahead with the largest ridge measure
                for (int a = -1; a < 2; a++)
                    for (int b = -1; b < 2; b++)
                        for (int c = -1; c < 2; c++) {
                            const int3 n(position.x + a, position.y + b, position.z + c);
                            const size_t ni = at(n.x, n.y, n.z);
                            if ((a == 0 && b == 0 && c == 0) || T.TDF[ni] == 0.0f) continue;
                            const float3 dir = normalized(float3((float)a, (float)b, (float)c));
                            if ((double)(dir.x * t_i.x + dir.y * t_i.y + dir.z * t_i.z) <= 0.1) continue;
                            const size_t mi = at(maxPoint.x, maxPoint.y, maxPoint.z);
                            if (T.radius[ni] >= 1.5f) {
                                if (M(ni) > M(mi)) maxPoint = n;
                            } else {
                                if (T.TDF[ni] * M(ni) > T.TDF[mi] * M(mi)) maxPoint = n;
                            }
                        }
                if (maxPoint.x + maxPoint.y + maxPoint.z <= 0) break;  // no candidate
                const size_t mi = at(maxPoint.x, maxPoint.y, maxPoint.z);
                if (centerlines[mi] > 0) {
                    // met an existing centerline
                    if (prevConnection == -1) {
                        prevConnection = centerlines[mi];
                    } else if (prevConnection == centerlines[mi]) {
                        connections = 5;  // a loop: reject this centerline
                    } else {
                        secondConnection = centerlines[mi];
                    }
                    break;
                }
                if (M(mi) < Mlow || (belowTlow > maxBelowTlow && T.TDF[mi] < Tlow)) break;
                if (newCenterlines.count(mi) > 0) break;  // loop within this centerline
                belowTlow = T.TDF[mi] < Tlow ? belowTlow + 1 : 0;
                // new direction: the eigenvector along the tube, kept pointing forward
                float3 lambda, e1, e2, e3;
                doEigen(T, maxPoint, size, &lambda, &e1, &e2, &e3);
                if (lambda.x < 0 && lambda.y < 0 && lambda.z < 0) {
                    if (std::fabs(dot3(t_i, e3)) > std::fabs(dot3(t_i, e2))) {
                        if (std::fabs(dot3(t_i, e3)) > std::fabs(dot3(t_i, e1))) e1 = e3;
                    } else if (std::fabs(dot3(t_i, e2)) > std::fabs(dot3(t_i, e1))) {
                        e1 = e2;
                    }
                }
                const float keep = sign_of(dot3(e1, t_i));
                const float3 sum(keep * e1.x + t_i.x + t_i_1.x, keep * e1.y + t_i.y + t_i_1.y,
                                 keep * e1.z + t_i.z + t_i_1.z);
                t_i_1 = t_i;
                t_i = normalized(sum);
                position = maxPoint;
                distance++;
                newCenterlines.insert(mi);
                meanTube += T.TDF[mi];
                CenterlinePoint cp;
                cp.pos = position;
                cp.next = &stack.top();
                cp.large = T.radius[mi] > 3.0f;
                stack.push(cp);
            }
        }

        if (distance > Dmin && meanTube / distance > minMeanTube && connections < 2) {
            if (prevConnection == -1) {
                for (size_t i : newCenterlines) centerlines[i] = counter;
                centerlineDistances[counter] = distance;
                centerlineStacks[counter] = stack;
                counter++;
            } else {
                std::stack<CenterlinePoint> prevStack = centerlineStacks[prevConnection];
                while (!stack.empty()) {
                    prevStack.push(stack.top());
                    stack.pop();
                }
                for (size_t i : newCenterlines) centerlines[i] = prevConnection;
                centerlineDistances[prevConnection] += distance;
                if (secondConnection != -1) {
                    // two connections: the second centerline joins the first
                    std::stack<CenterlinePoint> secondStack = centerlineStacks[secondConnection];
                    centerlineStacks.erase(secondConnection);
                    while (!secondStack.empty()) {
                        prevStack.push(secondStack.top());
                        secondStack.pop();
                    }
                    for (size_t i = 0; i < total; i++)
                        if (centerlines[i] == secondConnection) centerlines[i] = prevConnection;
                    centerlineDistances[prevConnection] += centerlineDistances[secondConnection];
                    centerlineDistances.erase(secondConnection);
                }
                centerlineStacks[prevConnection] = prevStack;
            }
        }
    }

    char* out = new char[total]();
    if (centerlineDistances.empty()) return out;
    // the longest tree, and every tree longer than min-tree-length
    int longest = centerlineDistances.begin()->first;
    std::list<int> trees;
    for (const auto& kv : centerlineDistances) {
        if (kv.second > centerlineDistances[longest]) longest = kv.first;
        if (kv.second > TreeMin) trees.push_back(kv.first);
    }
    centerlineStack = centerlineStacks[longest];  // a copy, as the reference's by-value argument
    std::unordered_set<int> keep(trees.begin(), trees.end());
    keep.insert(longest);
    for (size_t i = 0; i < total; i++) out[i] = keep.count(centerlines[i]) && centerlines[i] > 0 ? 1 : 0;
    return out;
}

void runCircleFittingAndRidgeTraversal(OpenCL& ocl, Image3D* dataset, int3* size, paramList& parameters,
                                       TSFOutput* output) {
    Image3D vectorField, radius;
    Image3D* TDF = new Image3D;
    runCircleFittingMethod(ocl, dataset, *size, parameters, vectorField, *TDF, radius);
    output->setTDF(TDF);
    // host arrays of the traversal: the field as float values (a 16-bit
    // field decoded as max(-1, s/32767)), the TDF (a 16-bit TDF already holds
    // t/65535) and the radius
    const size_t total = (size_t)size->x * size->y * size->z;
    TubeSegmentation TS;
    std::vector<float> fx = vectorField.downloadPlane(ocl, 0), fy = vectorField.downloadPlane(ocl, 1),
                       fz = vectorField.downloadPlane(ocl, 2), rad = radius.downloadPlane(ocl, 0);
    std::vector<float> tdf = TDF->downloadPlane(ocl, 0);
    TS.Fx = fx.data();
    TS.Fy = fy.data();
    TS.Fz = fz.data();
    TS.radius = rad.data();
    TS.TDF = new float[total];
    std::copy(tdf.begin(), tdf.end(), TS.TDF);
    output->setTDF(TS.TDF);
    char* centerline = runRidgeTraversal(TS, *size, parameters, std::stack<CenterlinePoint>());
    output->setCenterlineVoxels(centerline);
    if (!getParamBool(parameters, "no-segmentation")) {
        Buffer cl(total);
        cl.upload(ocl, centerline, total);
        Buffer* segmentation = new Buffer(
            getParamBool(parameters, "sphere-segmentation")
                ? runSphereSegmentation(ocl, cl, radius, *size, parameters)
                : runInverseGradientSegmentation(ocl, cl, vectorField, radius, *size, parameters));
        output->setSegmentation(segmentation);
    }
    if (getParamStr(parameters, "storage-dir") != "off")
        writeDataToDisk(output, getParamStr(parameters, "storage-dir"), getParamStr(parameters, "storage-name"));
}

}  // namespace tsf
