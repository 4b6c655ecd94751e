#include "griddingtaskmanager.h"

#include "hipdirectftgridder.h"
#include "hipwsmsgridder.h"

#include <stdexcept>

namespace wsclean_amd {

std::unique_ptr<GriddingTaskManager> GriddingTaskManager::Make(const Settings& settings) {
  if (settings.parallelGridding <= 1)
    return std::unique_ptr<GriddingTaskManager>(new GriddingTaskManager(settings));
  return std::unique_ptr<GriddingTaskManager>(new ThreadedScheduler(settings));
}

void GriddingTaskManager::Run(GriddingTask&& task,
                              std::function<void(GriddingResult&)> finishCallback) {
  GriddingResult result = RunDirect(std::move(task));
  finishCallback(result);
}

GriddingResult GriddingTaskManager::RunDirect(GriddingTask&& task) {
  std::unique_ptr<MSGridderBase> gridder = makeGridder(GetResources());
  return runDirect(std::move(task), *gridder);
}

GriddingResult GriddingTaskManager::runDirect(GriddingTask&& task, MSGridderBase& gridder) {
  gridder.ClearMeasurementSetList();
  for (MSProvider* p : task.msList) gridder.AddMeasurementSet(p);
  gridder.SetFacetGroupIndex(task.facetGroupIndex);
  gridder.SetIsFacet(false);
  gridder.SetImageWidth(settings_.paddedImageWidth);
  gridder.SetImageHeight(settings_.paddedImageHeight);
  gridder.SetTrimSize(settings_.trimmedImageWidth, settings_.trimmedImageHeight);
  gridder.SetLShift(task.l_shift);
  gridder.SetMShift(task.m_shift);
  // XY and YX are complex images (griddingtaskmanager.cpp:105-107)
  gridder.SetPolarization(task.complexPolarization ? aocommon::Polarization::XY
                                                   : task.polarization);
  gridder.SetIsComplex(task.complexPolarization);
  gridder.SetIsFirstIteration(task.verbose);
  if (task.cache)
    gridder.SetMetaDataCache(std::move(task.cache));
  else
    gridder.SetMetaDataCache(std::unique_ptr<MetaDataCache>(new MetaDataCache()));
  gridder.SetImageWeights(task.imageWeights.get());
  if (task.operation == GriddingTask::Invert) {
    gridder.SetPsfMode(task.imagePSF ? PsfMode::kSingle : PsfMode::kNone);
    gridder.SetDoSubtractModel(task.subtractModel);
    gridder.Invert();
  } else {
    gridder.SetWriterLockManager(this);
    gridder.Predict(std::move(task.modelImages));
  }
  GriddingResult result;
  result.images = gridder.ResultImages();
  result.startTime = gridder.StartTime();
  result.beamSize = gridder.BeamSize();
  result.imageWeight = gridder.ImageWeight();
  result.normalizationFactor = gridder.NormalizationFactor();
  result.actualWGridSize = gridder.ActualWGridSize();
  result.griddedVisibilityCount = gridder.GriddedVisibilityCount();
  result.effectiveGriddedVisibilityCount = gridder.EffectiveGriddedVisibilityCount();
  result.visibilityWeightSum = gridder.VisibilityWeightSum();
  result.cache = gridder.AcquireMetaDataCache();
  return result;
}

std::unique_ptr<MSGridderBase> GriddingTaskManager::constructGridder(
    const Resources& resources) const {
  switch (settings_.gridderType) {
    case GridderType::HipWStacking:
      return std::unique_ptr<MSGridderBase>(new ::HipWSMSGridder(settings_, resources));
    case GridderType::DirectFT:
      return std::unique_ptr<MSGridderBase>(new ::HipDirectFTGridder(settings_, resources));
    default:
      throw std::runtime_error(
          "only the MI355X w-stacking (hip-wstacking) and direct FT (direct-ft) gridders are "
          "available in this build");
  }
}

std::unique_ptr<MSGridderBase> GriddingTaskManager::makeGridder(const Resources& resources) const {
  std::unique_ptr<MSGridderBase> gridder = constructGridder(resources);
  gridder->SetGridMode(settings_.gridMode);
  return gridder;
}

// ---------------------------------------------------------------- threaded
ThreadedScheduler::ThreadedScheduler(const Settings& settings)
    : GriddingTaskManager(settings),
      resourcesPerTask_(GetResources().GetPart(settings.parallelGridding)) {}

ThreadedScheduler::~ThreadedScheduler() {
  if (!threads_.empty()) {
    try {
      Finish();
    } catch (std::exception&) {
      // a destructor may not throw; the caller's Finish() reports errors
    }
  }
}

void ThreadedScheduler::Run(GriddingTask&& task,
                            std::function<void(GriddingResult&)> finishCallback) {
  std::unique_lock<std::mutex> lock(mutex_);
  if (threads_.size() < settings_.parallelGridding) {
    threads_.emplace_back(&ThreadedScheduler::processQueue, this);
  } else {
    // block until a worker is free, so tasks do not pile up
    emptyCond_.wait(lock, [&] { return taskList_.empty(); });
  }
  while (!readyList_.empty()) {
    auto item = std::move(readyList_.back());
    readyList_.pop_back();
    lock.unlock();
    item.second(item.first);
    lock.lock();
  }
  taskList_.emplace_back(std::move(task), std::move(finishCallback));
  taskCond_.notify_one();
  lock.unlock();
  checkExceptions();
}

void ThreadedScheduler::processQueue() {
  std::unique_lock<std::mutex> lock(mutex_);
  for (;;) {
    taskCond_.wait(lock, [&] { return !taskList_.empty() || ended_; });
    if (taskList_.empty()) return;
    TaskPair taskPair = std::move(taskList_.front());
    taskList_.pop_front();
    if (taskList_.empty()) emptyCond_.notify_all();
    lock.unlock();
    try {
      std::unique_ptr<MSGridderBase> gridder = makeGridder(resourcesPerTask_);
      GriddingResult result = runDirect(std::move(taskPair.first), *gridder);
      lock.lock();
      readyList_.emplace_back(std::move(result), std::move(taskPair.second));
    } catch (std::exception&) {
      lock.lock();
      latestException_ = std::current_exception();
    }
  }
}

void ThreadedScheduler::Finish() {
  {
    std::lock_guard<std::mutex> lock(mutex_);
    ended_ = true;
    taskCond_.notify_all();
  }
  for (std::thread& t : threads_) t.join();
  threads_.clear();
  ended_ = false;
  while (!readyList_.empty()) {
    auto item = std::move(readyList_.back());
    readyList_.pop_back();
    item.second(item.first);
  }
  checkExceptions();
}

void ThreadedScheduler::checkExceptions() {
  std::exception_ptr toThrow;
  {
    std::lock_guard<std::mutex> lock(mutex_);
    toThrow = std::move(latestException_);
    latestException_ = std::exception_ptr();
  }
  if (toThrow) std::rethrow_exception(toThrow);
}

void ThreadedScheduler::Start(size_t nWriterGroups) {
  while (writerGroupLocks_.size() < nWriterGroups)
    writerGroupLocks_.emplace_back(new std::mutex());
}

std::unique_lock<std::mutex> ThreadedScheduler::GetLock(size_t writerGroupIndex) {
  if (writerGroupIndex >= writerGroupLocks_.size())
    throw std::runtime_error("writer group lock requested before Start()");
  return std::unique_lock<std::mutex>(*writerGroupLocks_[writerGroupIndex]);
}

}  // namespace wsclean_amd
