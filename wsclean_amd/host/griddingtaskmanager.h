// Gridding task flow of the reference: GriddingTask / GriddingResult
// (scheduling/griddingtask.h:23-59, griddingresult.h:13-42), the serial
// GriddingTaskManager (scheduling/griddingtaskmanager.{h,cpp}) and the
// ThreadedScheduler (scheduling/threadedscheduler.{h,cpp}). The MI355X
// gridder is registered in constructGridder as GridderType::HipWStacking;
// other gridder types are not part of this build and are rejected there.
#ifndef WSCLEAN_AMD_HOST_GRIDDINGTASKMANAGER_H
#define WSCLEAN_AMD_HOST_GRIDDINGTASKMANAGER_H

#include "msgridderbase.h"

#include <condition_variable>
#include <deque>
#include <exception>
#include <functional>
#include <memory>
#include <mutex>
#include <thread>
#include <vector>

namespace wsclean_amd {

struct GriddingTask {
  enum Operation { Invert, Predict } operation = Invert;
  bool imagePSF = false;
  bool subtractModel = false;
  bool complexPolarization = false;  // XY / YX (griddingtaskmanager.cpp:106-107)
  aocommon::PolarizationEnum polarization = aocommon::Polarization::StokesI;
  bool verbose = false;
  std::unique_ptr<MetaDataCache> cache;
  std::shared_ptr<ImageWeights> imageWeights;
  std::vector<MSProvider*> msList;  // providers of the task's correlation
  std::vector<Image> modelImages;
  double l_shift = 0.0, m_shift = 0.0;
  size_t facetGroupIndex = 0;
};

struct GriddingResult {
  std::vector<Image> images;
  double startTime = 0.0, beamSize = 0.0, imageWeight = 0.0, normalizationFactor = 0.0;
  size_t actualWGridSize = 0, griddedVisibilityCount = 0;
  double effectiveGriddedVisibilityCount = 0.0, visibilityWeightSum = 0.0;
  std::unique_ptr<MetaDataCache> cache;
};

class GriddingTaskManager : public WriterLockManager {
 public:
  virtual ~GriddingTaskManager() = default;
  // scheduling/griddingtaskmanager.cpp:24-39
  static std::unique_ptr<GriddingTaskManager> Make(const Settings& settings);
  virtual void Start(size_t /*nWriterGroups*/) {}
  std::unique_lock<std::mutex> GetLock(size_t) override {
    return std::unique_lock<std::mutex>(serialLock_);
  }
  // Runs the task now and calls the callback (:47-51).
  virtual void Run(GriddingTask&& task, std::function<void(GriddingResult&)> finishCallback);
  virtual void Finish() {}
  GriddingResult RunDirect(GriddingTask&& task);  // :53-56

 protected:
  explicit GriddingTaskManager(const Settings& settings) : settings_(settings) {}
  Resources GetResources() const {
    return Resources(settings_.threadCount, settings_.memory);
  }
  GriddingResult runDirect(GriddingTask&& task, MSGridderBase& gridder);  // :58-153
  std::unique_ptr<MSGridderBase> makeGridder(const Resources& resources) const;  // :182-187

  const Settings& settings_;

 private:
  std::unique_ptr<MSGridderBase> constructGridder(const Resources& resources) const;
  std::mutex serialLock_;
};

// -parallel-gridding N: N worker threads, each running whole tasks with its
// own gridder (own HIP context and stream) and Resources::GetPart(N).
// Callbacks run on the caller's thread; the latest exception of a task is
// rethrown at the next Run() or at Finish() (scheduling/threadedscheduler.cpp).
class ThreadedScheduler final : public GriddingTaskManager {
 public:
  explicit ThreadedScheduler(const Settings& settings);
  ~ThreadedScheduler() override;
  void Run(GriddingTask&& task, std::function<void(GriddingResult&)> finishCallback) override;
  void Finish() override;
  void Start(size_t nWriterGroups) override;
  std::unique_lock<std::mutex> GetLock(size_t writerGroupIndex) override;

 private:
  void processQueue();
  void checkExceptions();

  using TaskPair = std::pair<GriddingTask, std::function<void(GriddingResult&)>>;
  std::deque<TaskPair> taskList_;
  bool ended_ = false;
  std::condition_variable taskCond_, emptyCond_;
  std::mutex mutex_;
  std::vector<std::thread> threads_;
  std::vector<std::pair<GriddingResult, std::function<void(GriddingResult&)>>> readyList_;
  std::exception_ptr latestException_;
  const Resources resourcesPerTask_;
  std::vector<std::unique_ptr<std::mutex>> writerGroupLocks_;
  size_t busy_ = 0;
};

}  // namespace wsclean_amd

#endif
